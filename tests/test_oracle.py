"""Pins the CPU oracle (oracle/oracle.cpp) before it is trusted as the checker:
against the reference fixture's hand-derived answers, textbook routines,
closed forms, and with planted mistakes that the checks must catch."""
import json
import os

import numpy as np
import pytest

from starplat_amd import graph
from conftest import GOLDEN

EXP = json.load(open(os.path.join(GOLDEN, "sample_graph_expected.json")))


def cycle(n):
    s = np.arange(n, dtype=np.int32)
    return graph.from_edges(s, (s + 1) % n)


def complete(n):
    a, b = np.meshgrid(np.arange(n), np.arange(n))
    m = a != b
    return graph.from_edges(a[m].astype(np.int32), b[m].astype(np.int32))


def undirected(pairs, w=None):
    p = np.array(pairs, np.int32)
    s = np.concatenate([p[:, 0], p[:, 1]])
    d = np.concatenate([p[:, 1], p[:, 0]])
    ww = None if w is None else np.concatenate([w, w]).astype(np.int32)
    return graph.from_edges(s, d, ww)


def wheel(n):  # hub 0 + rim 1..n
    rim = [(i, i % n + 1) for i in range(1, n + 1)]
    return undirected([(0, i) for i in range(1, n + 1)] + rim)


def path(n):
    return undirected([(i, i + 1) for i in range(n - 1)])


def star(n):
    return undirected([(0, i) for i in range(1, n)])


# ---------------------------------------------------------------- golden
def test_golden_tc(oracle, sample_graph):
    assert oracle.compute_tc(sample_graph, literal=True) == EXP["triangles"]
    assert oracle.compute_tc(sample_graph, literal=False) == EXP["triangles"]


def test_golden_bc(oracle, sample_graph):
    bc, lev, sig = oracle.compute_bc(sample_graph, [0])
    np.testing.assert_array_equal(lev, EXP["bfs_levels_from_0"])
    np.testing.assert_array_equal(sig, EXP["sigma_from_0"])
    np.testing.assert_allclose(bc, EXP["bc_sources_0"], rtol=0, atol=1e-12)


def test_golden_sssp(oracle, sample_graph):
    d = oracle.compute_sssp(sample_graph, sample_graph.getEdgeLen(), 0)
    np.testing.assert_array_equal(d, EXP["sssp_from_0_zero_weights"])


# ---------------------------------------------------------------- SSSP
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_sssp_vs_dijkstra(oracle, seed):
    g = graph.rmat(11, 8, seed=seed, weights=(1, 100))
    for src in (0, 5, g.V - 1):
        np.testing.assert_array_equal(oracle.compute_sssp(g, g.edgeLen, src), oracle.dijkstra(g, g.edgeLen, src))


def test_sssp_closed_forms(oracle):
    gr = graph.grid2d(13, 17, weights=(1, 1))  # unit weights -> Manhattan distance
    d = oracle.compute_sssp(gr, gr.edgeLen, 0)
    r, c = np.divmod(np.arange(gr.V), 17)
    np.testing.assert_array_equal(d, r + c)
    # directed path with weights w: prefix sums; the reverse direction is unreachable
    w = np.array([3, 1, 4, 1, 5, 9, 2], np.int32)
    s = np.arange(8, dtype=np.int32)
    g = graph.from_edges(s[:-1], s[1:], w)
    np.testing.assert_array_equal(oracle.compute_sssp(g, g.edgeLen, 0), np.concatenate([[0], np.cumsum(w)]))
    d = oracle.compute_sssp(g, g.edgeLen, 3)
    assert np.all(d[:3] == np.iinfo(np.int32).max)


def test_sssp_planted_mistakes_caught(oracle):
    g = graph.rmat(11, 8, seed=9, weights=(1, 100))
    ref = oracle.dijkstra(g, g.edgeLen, 0)
    assert not np.array_equal(oracle.compute_sssp(g, g.edgeLen, 0, plant=2), ref)


# ---------------------------------------------------------------- PR
def test_pr_cycle_uniform(oracle):
    g = cycle(64)
    pr, it = oracle.compute_pr(g, beta=1e-10, delta=0.85, maxIter=100)
    # every vertex has in/out degree 1: ranks stay 1/n, diff = 0 after one iteration
    np.testing.assert_array_equal(pr, np.float32(1) / np.float32(64))
    assert it == 1


def test_pr_complete_uniform(oracle):
    g = complete(12)
    pr, _ = oracle.compute_pr(g)
    np.testing.assert_allclose(pr, 1 / 12, rtol=1e-6)


def test_pr_vs_fp64_power_iteration(oracle):
    g = graph.rmat(12, 16, seed=1)
    pr, it = oracle.compute_pr(g, 1e-10, 0.85, 100)
    ref, _ = oracle.pr_fp64(g, 1e-10, 0.85, it)
    assert np.max(np.abs(pr - ref)) / np.max(np.abs(ref)) < 1e-5


def test_pr_planted_mistakes_caught(oracle):
    g = graph.rmat(10, 8, seed=2)
    ref, _ = oracle.pr_fp64(g, 1e-10, 0.85, 100)
    for plant in (1, 2):
        with np.errstate(all="ignore"):
            pr, _ = oracle.compute_pr(g, plant=plant)
        assert not np.all(np.abs(pr - ref) <= 1e-5 * np.max(np.abs(ref)))
    # On a symmetric graph out-rows equal in-rows, so the wrong-array defect is
    # invisible there; a directed 3-cycle with a chord shows it.
    g2 = graph.from_edges(np.array([0, 1, 2, 0], np.int32), np.array([1, 2, 0, 2], np.int32))
    ok, _ = oracle.compute_pr(g2)
    ref2, _ = oracle.pr_fp64(g2)
    bad, _ = oracle.compute_pr(g2, plant=1)
    np.testing.assert_allclose(ok, ref2, rtol=1e-5)
    assert not np.allclose(bad, ref2, rtol=1e-5)


# ---------------------------------------------------------------- TC
@pytest.mark.parametrize("n", [3, 4, 7, 12])
def test_tc_complete(oracle, n):
    assert oracle.compute_tc(complete(n), literal=True) == n * (n - 1) * (n - 2) // 6


@pytest.mark.parametrize("n", [4, 5, 9])
def test_tc_wheel_grid_bipartite(oracle, n):
    assert oracle.compute_tc(wheel(n), literal=True) == n  # hub + n-cycle rim, n >= 4
    assert oracle.compute_tc(graph.grid2d(n, n + 2), literal=True) == 0
    bip = undirected([(i, n + j) for i in range(n) for j in range(n)])
    assert oracle.compute_tc(bip) == 0


def test_tc_vs_brute_and_merge(oracle):
    for seed in (1, 2, 3):
        g = graph.rmat(7, 8, seed=seed, undirected=True)
        t = oracle.tc_brute(g)
        assert oracle.compute_tc(g, literal=True) == t == oracle.tc_merge(g) == oracle.compute_tc(g)
    g = graph.rmat(14, 16, seed=3, undirected=True)
    assert oracle.compute_tc(g) == oracle.tc_merge(g)


def test_tc_planted_mistakes_caught(oracle):
    g = complete(6)
    for plant in (1, 2):
        assert oracle.compute_tc(g, plant=plant) != 20


# ---------------------------------------------------------------- BC
def test_bc_path_closed_form(oracle):
    n = 9
    bc, lev, sig = oracle.compute_bc(path(n), [0])
    # from source 0 on a path, vertex v has n-1-v descendants
    np.testing.assert_allclose(bc, [0] + [n - 1 - v for v in range(1, n)], atol=0)
    np.testing.assert_array_equal(lev, np.arange(n))
    np.testing.assert_array_equal(sig, 1)


def test_bc_star_closed_form(oracle):
    n = 8
    leaves = list(range(1, n))
    bc, _, _ = oracle.compute_bc(star(n), leaves)
    # centre lies on every leaf-to-other-leaf path: (n-2) per leaf source
    assert bc[0] == (n - 1) * (n - 2)
    np.testing.assert_allclose(bc[1:], 0)


def test_bc_dependency_sum_invariant(oracle):
    g = graph.rmat(11, 8, seed=4, undirected=True)
    for s in (0, 17, 300):
        bc, lev, _ = oracle.compute_bc(g, [s])
        reach = lev >= 0
        assert bc.sum() == pytest.approx(float(np.sum(lev[reach] - 1) + 1), rel=1e-9)


@pytest.mark.parametrize("seed", [1, 2])
def test_bc_vs_textbook_brandes(oracle, seed):
    g = graph.rmat(11, 8, seed=seed)  # directed
    srcs = [0, 3, 77, 1000]
    bc, lev, sig = oracle.compute_bc(g, srcs)
    tb, tlev, tsig = oracle.brandes(g, srcs)
    np.testing.assert_array_equal(lev, tlev)
    np.testing.assert_array_equal(sig, tsig)
    np.testing.assert_allclose(bc, tb, rtol=1e-12, atol=1e-12)


def test_bc_planted_mistakes_caught(oracle, sample_graph):
    for plant in (1, 2):
        bc, _, sig = oracle.compute_bc(sample_graph, [0], plant=plant)
        assert not (np.allclose(bc, EXP["bc_sources_0"]) and np.array_equal(sig, EXP["sigma_from_0"]))
