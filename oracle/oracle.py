"""TEST INFRASTRUCTURE ONLY: ctypes wrapper of liboracle.so, the CPU
restatement of StarPlat's generated OpenMP code (oracle.cpp) and the textbook
routines that pin it. Imported by the test suite, the smoke check and the CPU
baseline leg of the benchmark; never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

i32, i64, f32, f64 = C.c_int32, C.c_int64, C.c_float, C.c_double
P_i32, P_f32, P_f64 = C.POINTER(i32), C.POINTER(f32), C.POINTER(f64)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing: run `make -C {_HERE}`")
        L = C.CDLL(LIB)
        L.oracle_compute_pr.argtypes = [i32, P_i32, P_i32, P_i32, P_i32, f32, f32, i32, P_f32, P_i32, C.c_int]
        L.oracle_compute_sssp.argtypes = [i32, P_i32, P_i32, P_i32, P_i32, i32, C.c_int]
        L.oracle_compute_tc.argtypes = [i32, P_i32, P_i32, C.c_int, C.c_int]
        L.oracle_compute_tc.restype = i64
        L.oracle_compute_tc_strided.argtypes = [i32, P_i32, P_i32, i32]
        L.oracle_compute_tc_strided.restype = i64
        L.oracle_compute_bc.argtypes = [i32, P_i32, P_i32, P_i32, i32, P_f64, P_i32, P_f64, C.c_int]
        L.tb_dijkstra.argtypes = [i32, P_i32, P_i32, P_i32, i32, P_i32]
        L.tb_brandes.argtypes = [i32, P_i32, P_i32, P_i32, i32, P_f64, P_i32, P_f64]
        L.tb_tc_brute.argtypes = [i32, P_i32, P_i32]
        L.tb_tc_brute.restype = i64
        L.tb_tc_merge.argtypes = [i32, P_i32, P_i32]
        L.tb_tc_merge.restype = i64
        L.tb_pr_fp64.argtypes = [i32, P_i32, P_i32, P_i32, f64, f64, i32, P_f64, P_i32]
        L.oracle_dyn_sssp.argtypes = [i32, P_i32, P_i32, P_i32, i32, i32, C.c_char_p, P_i32, P_i32, i32, P_i32,
                                      P_i32, C.c_int, P_f64]
        L.oracle_dyn_sssp.restype = C.c_int
        L.oracle_dyn_tc.argtypes = [i32, P_i32, P_i32, i32, C.c_char_p, P_i32, P_i32, i32, C.c_int64,
                                    C.POINTER(C.c_int64), C.c_int, P_f64]
        L.oracle_dyn_tc.restype = C.c_int32
        L.oracle_dyn_pr.argtypes = [i32, P_i32, P_i32, C.c_float, C.c_float, i32, i32, C.c_char_p, P_i32, P_i32, i32,
                                    C.POINTER(C.c_float), P_i32, C.c_int, P_f64]
        L.oracle_dyn_pr.restype = C.c_int32
        L.oracle_num_threads.restype = C.c_int
        _lib = L
    return _lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def num_threads():
    return lib().oracle_num_threads()


def compute_pr(g, beta=1e-10, delta=0.85, maxIter=100, plant=0):
    pr = np.empty(g.V, np.float32)
    it = i32(0)
    lib().oracle_compute_pr(g.V, _p(g.indexofNodes, i32), _p(g.edgeList, i32), _p(g.rev_indexofNodes, i32),
                            _p(g.srcList, i32), beta, delta, maxIter, _p(pr, f32), C.byref(it), plant)
    return pr, it.value


def compute_sssp(g, weight, src, plant=0):
    d = np.empty(g.V, np.int32)
    w = np.ascontiguousarray(weight, np.int32)
    lib().oracle_compute_sssp(g.V, _p(g.indexofNodes, i32), _p(g.edgeList, i32), _p(d, i32), _p(w, i32), src, plant)
    return d


def compute_tc(g, literal=False, plant=0):
    return int(lib().oracle_compute_tc(g.V, _p(g.indexofNodes, i32), _p(g.edgeList, i32), int(literal), plant))


def compute_tc_strided(g, stride):
    return int(lib().oracle_compute_tc_strided(g.V, _p(g.indexofNodes, i32), _p(g.edgeList, i32), stride))


def compute_bc(g, sources, plant=0):
    """FP64 BC plus the bfsDist / sigma of the last source."""
    s = np.ascontiguousarray(sorted(set(int(x) for x in sources)), np.int32)
    bc = np.empty(g.V, np.float64)
    lev = np.empty(g.V, np.int32)
    sig = np.empty(g.V, np.float64)
    lib().oracle_compute_bc(g.V, _p(g.indexofNodes, i32), _p(g.edgeList, i32), _p(s, i32), len(s),
                            _p(bc, f64), _p(lev, i32), _p(sig, f64), plant)
    return bc, lev, sig


def dijkstra(g, weight, src):
    d = np.empty(g.V, np.int32)
    w = np.ascontiguousarray(weight, np.int32)
    lib().tb_dijkstra(g.V, _p(g.indexofNodes, i32), _p(g.edgeList, i32), _p(w, i32), src, _p(d, i32))
    return d


def brandes(g, sources):
    s = np.ascontiguousarray(sorted(set(int(x) for x in sources)), np.int32)
    bc = np.empty(g.V, np.float64)
    lev = np.empty(g.V, np.int32)
    sig = np.empty(g.V, np.float64)
    lib().tb_brandes(g.V, _p(g.indexofNodes, i32), _p(g.edgeList, i32), _p(s, i32), len(s), _p(bc, f64),
                     _p(lev, i32), _p(sig, f64))
    return bc, lev, sig


def tc_brute(g):
    return int(lib().tb_tc_brute(g.V, _p(g.indexofNodes, i32), _p(g.edgeList, i32)))


def tc_merge(g):
    return int(lib().tb_tc_merge(g.V, _p(g.indexofNodes, i32), _p(g.edgeList, i32)))


def pr_fp64(g, beta=1e-10, delta=0.85, maxIter=100):
    pr = np.empty(g.V, np.float64)
    it = i32(0)
    lib().tb_pr_fp64(g.V, _p(g.indexofNodes, i32), _p(g.rev_indexofNodes, i32), _p(g.srcList, i32), beta, delta,
                     maxIter, _p(pr, f64), C.byref(it))
    return pr, it.value


def dyn_sssp(g, weight, src, types, us, ut, batch, plant=0, timing=None):
    """DynSSSP restated (oracle.cpp): (dist, parent) after every batch of
    `batch` updates; types is a bytes string of 'a' / 'd'. With a dict as
    `timing`, timing["batch_seconds"] receives the batch loop's wall time."""
    w = np.ascontiguousarray(weight, np.int32)
    us = np.ascontiguousarray(us, np.int32)
    ut = np.ascontiguousarray(ut, np.int32)
    t = bytes(types)
    assert len(t) == len(us) == len(ut)
    dist = np.empty(g.V, np.int32)
    parent = np.empty(g.V, np.int32)
    secs = f64(0.0)
    rc = lib().oracle_dyn_sssp(g.V, _p(g.indexofNodes, i32), _p(g.edgeList, i32), _p(w, i32), int(src), len(us), t,
                               _p(us, i32), _p(ut, i32), int(batch), _p(dist, i32), _p(parent, i32), int(plant),
                               C.byref(secs))
    if timing is not None:
        timing["batch_seconds"] = secs.value
    if rc < 0:
        raise RuntimeError("oracle_dyn_sssp: a loop did not converge")
    return dist, parent


def dyn_tc(g, types, us, ut, batch, plant=0, timing=None, initial=None):
    """DynTC restated (oracle.cpp): the triangle count after every batch of
    `batch` updates; types is a bytes string of 'a' / 'd'. With a dict as
    `timing`, timing["batch_seconds"] receives the batch loop's wall time.
    `initial` replaces the static count the run starts from (for timing)."""
    us = np.ascontiguousarray(us, np.int32)
    ut = np.ascontiguousarray(ut, np.int32)
    t = bytes(types)
    assert len(t) == len(us) == len(ut)
    nb = (len(us) + batch - 1) // batch
    counts = np.zeros(max(nb, 1), np.int64)
    secs = f64(0.0)
    got = lib().oracle_dyn_tc(g.V, _p(g.indexofNodes, i32), _p(g.edgeList, i32), len(us), t, _p(us, i32),
                              _p(ut, i32), int(batch), -1 if initial is None else int(initial),
                              _p(counts, C.c_int64), int(plant), C.byref(secs))
    if timing is not None:
        timing["batch_seconds"] = secs.value
    assert got == nb
    return counts[:nb]


def dyn_pr(g, beta, delta, maxIter, types, us, ut, batch, plant=0, timing=None):
    """DynPR restated (oracle.cpp): (ranks after the last batch, iteration
    counts: staticPR's, then Decremental's and Incremental's per batch)."""
    us = np.ascontiguousarray(us, np.int32)
    ut = np.ascontiguousarray(ut, np.int32)
    t = bytes(types)
    assert len(t) == len(us) == len(ut)
    nb = (len(us) + batch - 1) // batch
    pr = np.empty(g.V, np.float32)
    its = np.zeros(1 + 2 * nb, np.int32)
    secs = f64(0.0)
    got = lib().oracle_dyn_pr(g.V, _p(g.indexofNodes, i32), _p(g.edgeList, i32), float(beta), float(delta),
                              int(maxIter), len(us), t, _p(us, i32), _p(ut, i32), int(batch),
                              pr.ctypes.data_as(C.POINTER(C.c_float)), _p(its, i32), int(plant), C.byref(secs))
    if timing is not None:
        timing["batch_seconds"] = secs.value
    assert got == nb
    return pr, its


def apply_updates(V, off, adj, w, ops, plant=0):
    """Edge-multiset restatement of updateCSRDel / updateCSRAdd
    (graph.hpp:869-975): ops is a sequence of ("d", src, dst) or
    ("a", src, dst, weight) array tuples applied in order; a deletion removes
    every copy of the edge, an addition appends one. Returns the CSR with rows
    sorted by (neighbour, weight): (indexofNodes, edgeList, edgeLen).
    plant 1: a deletion removes only the first copy."""
    off = np.asarray(off, np.int64)
    s = np.repeat(np.arange(V, dtype=np.int64), np.diff(off))
    d = np.asarray(adj, np.int64).copy()
    ww = np.asarray(w, np.int64).copy()
    for op in ops:
        if op[0] == "d":
            for a, b in zip(np.asarray(op[1], np.int64), np.asarray(op[2], np.int64)):
                hit = (s == a) & (d == b)
                if plant == 1 and hit.any():
                    first = np.argmax(hit)
                    hit = np.zeros_like(hit)
                    hit[first] = True
                keep = ~hit
                s, d, ww = s[keep], d[keep], ww[keep]
        else:
            s = np.concatenate([s, np.asarray(op[1], np.int64)])
            d = np.concatenate([d, np.asarray(op[2], np.int64)])
            ww = np.concatenate([ww, np.asarray(op[3], np.int64)])
    order = np.lexsort((ww, d, s))
    s, d, ww = s[order], d[order], ww[order]
    new_off = np.zeros(V + 1, np.int64)
    np.add.at(new_off, s + 1, 1)
    return np.cumsum(new_off).astype(np.int32), d.astype(np.int32), ww.astype(np.int32)
