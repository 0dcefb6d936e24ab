"""The four generated entry points, with the reference's names and argument
meaning, running on the gfx950 kernel library through the C ABI.

    Compute_PR(g, beta, delta, maxIter, pageRank)   PageRankDSLV2 / V3 diff
    Compute_SSSP(g, dist, weight, src)              sssp_dslV2
    Compute_TC(g) -> int                            triangle_counting_dsl
    Compute_BC(g, BC, sourceSet)                    bc_dslV2
    DynSSSP(g, dist, parent, weight, updateBatch, batchSize, src)
                                                    dynamicBatchSSSP

Output arrays are caller-owned numpy arrays of num_nodes() entries, filled in
place, as the generated C++ fills the caller's pointers
(backend_omp/dsl_cpp_generator.cpp:3056-3151). A failing call raises
``SpmError``; nothing falls back to the CPU.
"""
from __future__ import annotations

import numpy as np


def _out(a, dtype, n, name):
    if not isinstance(a, np.ndarray) or a.dtype != dtype or a.size != n or not a.flags.c_contiguous:
        raise TypeError(f"{name} must be a contiguous numpy {np.dtype(dtype).name} array of num_nodes() entries")
    return a


def Compute_PR(g, beta, delta, maxIter, pageRank):
    """Pull PageRank; returns the number of iterations run."""
    _out(pageRank, np.float32, g.num_nodes(), "pageRank")
    _, it = g.device().pr(float(beta), float(delta), int(maxIter), out=pageRank)
    return it


def Compute_SSSP(g, dist, weight, src):
    """Bellman-Ford fixed point; dist[v] = INT_MAX when unreachable."""
    _out(dist, np.int32, g.num_nodes(), "dist")
    w = np.ascontiguousarray(weight, np.int32)
    if w.size != g.num_edges():
        raise ValueError("weight needs one entry per CSR edge (indexed by edge position)")
    g.device().sssp(int(src), weight=w, out=dist)


def Compute_TC(g):
    """Number of (u < v < w) wedges closed by an edge, as the DSL sums them."""
    return int(g.device().tc())


def Compute_BC(g, BC, sourceSet):
    """Brandes BC accumulated over sourceSet (no 1/2 factor, as the DSL)."""
    _out(BC, np.float32, g.num_nodes(), "BC")
    src = np.array(sorted(set(int(s) for s in sourceSet)), np.int32)
    g.device().bc(src, out=BC)


def DynSSSP(g, dist, parent, weight, updateBatch, batchSize, src):
    """The Dynamic function of graphcode/dynamicDSLCodes/dynamicBatchSSSP:67-94:
    staticSSSP from src, then per batch of batchSize updates OnDelete,
    updateCSRDel, Decremental, OnAdd, updateCSRAdd, Incremental (added edges
    weigh 1). weight: one entry per CSR edge of g before the updates (None
    keeps the device mirror's). updateBatch: a list of ``updates.update``.
    dist / parent are filled in place; g's arrays then describe the updated
    graph, as the reference mutates its graph."""
    from .updates import as_arrays
    V = g.num_nodes()
    _out(dist, np.int32, V, "dist")
    _out(parent, np.int32, V, "parent")
    dev = g.device()
    if weight is not None:
        w = np.ascontiguousarray(weight, np.int32)
        if w.size != dev.num_edges():
            raise ValueError("weight needs one entry per CSR edge (indexed by edge position)")
        dev.set_weights(w)
    types, s, d, _ = as_arrays(list(updateBatch))
    dd, pp = dev.dyn_sssp(int(src), types, s, d, int(batchSize))
    dist[:] = dd
    parent[:] = pp
    g.refresh_from_device()


def DynTC(g, updateBatch, batchSize):
    """The Dynamic function of graphcode/dynamicDSLCodes/dynamicBatchTCV2:137-187:
    staticTC, then per batch of batchSize updates OnDelete + Decremental on the
    graph before the deletions, updateCSRDel, updateCSRAdd, OnAdd +
    Incremental. updateBatch: a list of ``updates.update``. g's arrays then
    describe the updated graph. The DSL function returns nothing; the
    triangle count it keeps is returned here."""
    from .updates import as_arrays
    dev = g.device()
    types, s, d, _ = as_arrays(list(updateBatch))
    count = dev.dyn_tc(types, s, d, int(batchSize))
    g.refresh_from_device()
    return count


def DynPR(g, beta, delta, maxIter, pageRank, updateBatch, batchSize):
    """The Dynamic function of graphcode/dynamicDSLCodes/dynamicBatchPR:97-135
    as the dynamic OMP generator emits it: staticPR, then per batch of
    batchSize updates the flagged sweeps Decremental (after updateCSRDel) and
    Incremental (after updateCSRAdd); count_outNbrs is the out-degree of the
    graph the run starts from and the diff is signed. pageRank is filled in
    place; g's arrays then describe the updated graph."""
    from .updates import as_arrays
    V = g.num_nodes()
    _out(pageRank, np.float32, V, "pageRank")
    dev = g.device()
    types, s, d, _ = as_arrays(list(updateBatch))
    pageRank[:] = dev.dyn_pr(float(beta), float(delta), int(maxIter), types, s, d, int(batchSize))
    g.refresh_from_device()
