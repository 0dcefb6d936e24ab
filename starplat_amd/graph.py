"""``graph``: host mirror of StarPlat's CSR graph class (graphcode/graph.hpp).

Same public arrays and methods the generated code uses:

    indexofNodes      graph.hpp:44      edgeList     :46
    rev_indexofNodes  :45               srcList      :48
    getEdgeLen()      :134-137          num_nodes()  :139-142   num_edges() :145-152
    parseGraph()      :861-865          check_if_nbr :233-287   getEdge     :168-186

The arrays are numpy int32 views of memory owned by the native host library;
the CSR is built by a parallel sort rather than graph.hpp's std::map, with the
same contents. Seeded generators (RMAT, Erdos-Renyi, grid) produce graphs
directly. The device mirror (an ``spm_graph`` handle holding the CSR in HBM)
is created on first use by the Compute_* functions and reused after.
"""
from __future__ import annotations

import ctypes as C
import weakref

import numpy as np

from . import _native as N


def _view(p, n):
    if n == 0:
        return np.zeros(0, np.int32)
    return np.ctypeslib.as_array(p, shape=(int(n),))


class graph:
    """CSR graph with graph.hpp semantics (num_nodes = max id + 1)."""

    def __init__(self, filePath: str | None = None):
        self.filePath = filePath
        self._h = None
        self._dev = None
        self._dev_weights = None

    # ---- construction --------------------------------------------------
    def _adopt(self, h):
        if not h:
            raise RuntimeError("graph construction failed (bad arguments or unreadable file)")
        self._h = h
        hl = N.host_lib()
        weakref.finalize(self, hl.spmh_free, h)
        s = h.contents
        self.V = int(s.V)
        self.E = int(s.E)
        self.indexofNodes = _view(s.indexofNodes, self.V + 1)
        self.edgeList = _view(s.edgeList, self.E)
        self.edgeLen = _view(s.edgeLen, self.E)
        self.rev_indexofNodes = _view(s.rev_indexofNodes, self.V + 1)
        self.srcList = _view(s.srcList, self.E)
        return self

    def _native(self):
        """The native host structure (spmh_graph*) behind the arrays."""
        if self._h is None:
            raise RuntimeError("the graph's native host structure is stale after edge updates; "
                               "statistics and writers need a graph built from edges or a file")
        return self._h

    def parseGraph(self):
        """graph::parseGraph(): parseEdges (graph.hpp:464-515) + parseEdgesContent (:720-859)."""
        if self.filePath is None:
            raise ValueError("graph has no file path")
        return self._adopt(N.host_lib().spmh_parse_text(str(self.filePath).encode()))

    @classmethod
    def rmat(cls, scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1, undirected=False,
             weights=None):
        """RMAT graph: edge_factor*2^scale samples, self-loops and duplicates removed.
        weights=(lo, hi) assigns uniform int32 weights per surviving edge."""
        lo, hi = weights if weights else (0, 0)
        return cls()._adopt(N.host_lib().spmh_rmat(scale, edge_factor, a, b, c, seed,
                                                   int(bool(undirected)), lo, hi))

    @classmethod
    def erdos_renyi(cls, n, m, seed=1, undirected=True, weights=None):
        lo, hi = weights if weights else (0, 0)
        return cls()._adopt(N.host_lib().spmh_er(n, m, seed, int(bool(undirected)), lo, hi))

    @classmethod
    def grid2d(cls, rows, cols, seed=1, weights=None):
        lo, hi = weights if weights else (0, 0)
        return cls()._adopt(N.host_lib().spmh_grid(rows, cols, seed, lo, hi))

    @classmethod
    def from_edges(cls, src, dst, weight=None):
        """graph.hpp semantics on an explicit edge list: duplicates summed, self-loops kept."""
        src = np.ascontiguousarray(src, np.int32)
        dst = np.ascontiguousarray(dst, np.int32)
        w = None if weight is None else np.ascontiguousarray(weight, np.int32)
        if src.shape != dst.shape or (w is not None and w.shape != src.shape):
            raise ValueError("src, dst and weight must have the same length")
        return cls()._adopt(N.host_lib().spmh_from_edges(
            len(src), N.ptr(src, C.c_int32), N.ptr(dst, C.c_int32), N.ptr(w, C.c_int32)))

    @classmethod
    def union(cls, graphs):
        """Disjoint union; returns (graph, bases) with graph i's vertex v at bases[i] + v."""
        arr = (N.P_spmh * len(graphs))(*[g._native() for g in graphs])
        bases = np.zeros(len(graphs) + 1, np.int64)
        u = cls()._adopt(N.host_lib().spmh_union(arr, len(graphs), N.ptr(bases, C.c_int64)))
        return u, bases

    def save(self, path):
        """Binary CSR file (header + the five arrays, checksummed)."""
        if N.host_lib().spmh_write_binary(self._native(), str(path).encode()) != 0:
            raise OSError(f"cannot write {path}")

    @classmethod
    def load(cls, path):
        h = N.host_lib().spmh_read_binary(str(path).encode())
        if not h:
            raise OSError(f"{path}: missing, truncated or corrupt binary CSR")
        return cls()._adopt(h)

    def write_text(self, path, weights=True):
        rc = N.host_lib().spmh_write_text(self._native(), str(path).encode(), int(bool(weights)))
        if rc != 0:
            raise OSError(f"cannot write {path}")

    # ---- graph.hpp accessors -------------------------------------------
    def num_nodes(self):
        return self.V

    def num_edges(self):
        return self.E

    def getEdgeLen(self):
        return self.edgeLen

    @property
    def edgeMap(self):
        """graph.hpp's edgeMap (:57, :835-846): CSR position of the edge at
        every reverse-CSR slot."""
        if getattr(self, "_edge_map", None) is None or self._edge_map.size != self.E:
            m = np.empty(self.E, np.int32)
            rc = N.host_lib().spmh_edge_map(self.V, N.ptr(self.indexofNodes, C.c_int32), N.ptr(self.edgeList, C.c_int32),
                                            N.ptr(self.rev_indexofNodes, C.c_int32), N.ptr(self.srcList, C.c_int32),
                                            N.ptr(m, C.c_int32))
            if rc != 0:
                raise RuntimeError("edgeMap: the CSR and the reverse CSR disagree")
            self._edge_map = m
        return self._edge_map

    def getOutDegree(self, v):
        """graph.hpp:316-320."""
        return int(self.indexofNodes[v + 1] - self.indexofNodes[v])

    def getInDegree(self, v):
        """graph.hpp:322-326."""
        return int(self.rev_indexofNodes[v + 1] - self.rev_indexofNodes[v])

    def count_outNbrs(self, v):
        return int(self.indexofNodes[v + 1] - self.indexofNodes[v])

    def neighbors(self, v):
        return self.edgeList[self.indexofNodes[v]:self.indexofNodes[v + 1]]

    def nodes_to(self, v):
        return self.srcList[self.rev_indexofNodes[v]:self.rev_indexofNodes[v + 1]]

    def check_if_nbr(self, s, d):
        return bool(np.any(self.neighbors(s) == d))

    def getEdge(self, s, d):
        """(weight, CSR position) of the first edge s->d, as getEdge returns it."""
        row = self.neighbors(s)
        hit = np.nonzero(row == d)[0]
        if len(hit) == 0:
            return None
        e = int(self.indexofNodes[s] + hit[0])
        return int(self.edgeLen[e]), e

    def is_symmetric(self):
        fwd = np.repeat(np.arange(self.V, dtype=np.int64), np.diff(self.indexofNodes)) * (1 << 32) + self.edgeList
        rev = self.edgeList.astype(np.int64) * (1 << 32) + np.repeat(np.arange(self.V, dtype=np.int64), np.diff(self.indexofNodes))
        return np.array_equal(np.sort(fwd), np.sort(rev))

    # ---- device mirror ---------------------------------------------------
    def device(self):
        """The spm_graph handle mirroring this CSR in HBM (created once)."""
        if self._dev is None:
            from .device import DeviceGraph
            self._dev = DeviceGraph(self)
        return self._dev

    def release_device(self):
        if self._dev is not None:
            self._dev.close()
            self._dev = None

    # ---- edge updates (graph.hpp:869-975) -------------------------------
    def updateCSRDel(self, batchUpdate, k, size):
        """graph::updateCSRDel (graph.hpp:869-930): the deletions among
        batchUpdate[k:k+size], applied to the device mirror; the host arrays
        are then refreshed from it."""
        from .updates import as_arrays
        dels = [u for u in batchUpdate[k:k + size] if u.type == "d"]
        if dels:
            _, s, d, _ = as_arrays(dels)
            self.device().delete_edges(s, d)
        self.refresh_from_device()

    def updateCSRAdd(self, batchUpdate, k, size):
        """graph::updateCSRAdd (graph.hpp:932-975): the additions among
        batchUpdate[k:k+size] with their weights."""
        from .updates import as_arrays
        adds = [u for u in batchUpdate[k:k + size] if u.type == "a"]
        if adds:
            _, s, d, w = as_arrays(adds)
            self.device().add_edges(s, d, w)
        self.refresh_from_device()

    def refresh_from_device(self):
        """Replaces the host arrays with the device mirror's current CSR (rows
        ascending). The native host structure no longer describes the graph
        afterwards, so generators' statistics (features, byte models) refuse
        it."""
        off, adj, w, roff, radj = self.device().export()
        self.indexofNodes, self.edgeList = off, adj
        self.E = int(adj.size)
        self.edgeLen = w if w is not None else np.zeros(self.E, np.int32)
        if roff is not None:
            self.rev_indexofNodes, self.srcList = roff, radj
        self._edge_map = None
        self._h = None


def sources(V, n, seed):
    """n distinct uniform vertices of [0, V), ascending (a std::set<int>'s order)."""
    out = np.empty(n, np.int32)
    rc = N.host_lib().spmh_sources(V, n, seed, N.ptr(out, C.c_int32))
    if rc != 0:
        raise ValueError("cannot draw that many distinct sources")
    return out
