"""DeviceGraph: Python owner of an ``spm_graph`` handle (include/starplat_amd.h).

Wraps each C-ABI entry point with numpy buffers and raises ``SpmError`` on a
non-zero status. Methods whose name ends in ``_device`` take and fill
``hip.DeviceArray`` outputs and leave results in HBM.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N


class DeviceGraph:
    def __init__(self, g):
        self.V = g.num_nodes()
        self.E = g.num_edges()
        lib = N.device_lib()
        h = C.c_void_p()
        w = g.edgeLen if g.E else None
        rc = lib.spm_graph_create(
            self.V, self.E, N.ptr(g.indexofNodes, C.c_int32), N.ptr(g.edgeList, C.c_int32),
            N.ptr(w, C.c_int32), N.ptr(g.rev_indexofNodes, C.c_int32), N.ptr(g.srcList, C.c_int32),
            C.byref(h))
        N.check("spm_graph_create", rc)
        self.h = h.value
        self.has_weights = w is not None
        self.has_reverse = g.rev_indexofNodes is not None and g.srcList is not None

    def close(self):
        if self.h:
            N.device_lib().spm_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self):
        return N.device_lib().spm_graph_stream(self.h)

    def device_bytes(self):
        return int(N.device_lib().spm_graph_device_bytes(self.h))

    def set_weights(self, weight):
        weight = np.ascontiguousarray(weight, np.int32)
        self.num_edges()
        if weight.size != self.E:
            raise ValueError("weight must have one entry per CSR edge")
        N.check("spm_graph_set_weights", N.device_lib().spm_graph_set_weights(self.h, N.ptr(weight, C.c_int32)))
        self.has_weights = True

    # ---- edge updates -------------------------------------------------------
    @staticmethod
    def _pair(src, dst):
        src = np.ascontiguousarray(src, np.int32)
        dst = np.ascontiguousarray(dst, np.int32)
        if src.shape != dst.shape or src.ndim != 1:
            raise ValueError("src and dst must be 1-D arrays of one length")
        return src, dst

    def delete_edges(self, src, dst):
        """graph::updateCSRDel for the edges (src[i], dst[i])."""
        src, dst = self._pair(src, dst)
        N.check("spm_graph_delete_edges", N.device_lib().spm_graph_delete_edges(
            self.h, N.ptr(src, C.c_int32), N.ptr(dst, C.c_int32), len(src)))

    def add_edges(self, src, dst, weight=None):
        """graph::updateCSRAdd for the edges (src[i], dst[i]) (weight 1 when None)."""
        src, dst = self._pair(src, dst)
        if weight is not None:
            weight = np.ascontiguousarray(weight, np.int32)
            if weight.shape != src.shape:
                raise ValueError("weight must match src")
        N.check("spm_graph_add_edges", N.device_lib().spm_graph_add_edges(
            self.h, N.ptr(src, C.c_int32), N.ptr(dst, C.c_int32), N.ptr(weight, C.c_int32), len(src)))

    def compact(self):
        N.check("spm_graph_compact", N.device_lib().spm_graph_compact(self.h))

    def num_edges(self):
        e = C.c_int64(0)
        N.check("spm_graph_num_edges", N.device_lib().spm_graph_num_edges(self.h, C.byref(e)))
        self.E = int(e.value)
        return self.E

    def export(self):
        """The current CSR (after folding any updates) as numpy arrays:
        (indexofNodes, edgeList, edgeLen or None, rev_indexofNodes or None,
        srcList or None)."""
        E = self.num_edges()
        off = np.empty(self.V + 1, np.int32)
        adj = np.empty(E, np.int32)
        w = np.empty(E, np.int32) if self.has_weights else None
        roff = np.empty(self.V + 1, np.int32) if self.has_reverse else None
        radj = np.empty(E, np.int32) if self.has_reverse else None
        N.check("spm_graph_export", N.device_lib().spm_graph_export(
            self.h, N.ptr(off, C.c_int32), N.ptr(adj, C.c_int32), N.ptr(w, C.c_int32),
            N.ptr(roff, C.c_int32), N.ptr(radj, C.c_int32)))
        return off, adj, w, roff, radj

    # ---- dynamic SSSP ---------------------------------------------------------
    @staticmethod
    def _types(types, n):
        t = types.encode() if isinstance(types, str) else bytes(types)
        if len(t) != n:
            raise ValueError("one update type per update")
        return t

    def dynsssp_init(self, src):
        N.check("spm_dynsssp_init", N.device_lib().spm_dynsssp_init(self.h, int(src)))

    def dynsssp_batch(self, types, src, dst):
        src, dst = self._pair(src, dst)
        t = self._types(types, len(src))
        N.check("spm_dynsssp_batch", N.device_lib().spm_dynsssp_batch(
            self.h, t, N.ptr(src, C.c_int32), N.ptr(dst, C.c_int32), len(src)))

    def dynsssp_result(self):
        dist = np.empty(self.V, np.int32)
        parent = np.empty(self.V, np.int32)
        N.check("spm_dynsssp_result", N.device_lib().spm_dynsssp_result(
            self.h, N.ptr(dist, C.c_int32), N.ptr(parent, C.c_int32)))
        return dist, parent

    def dynsssp_device(self, d_dist, d_parent=None):
        N.check("spm_dynsssp_device", N.device_lib().spm_dynsssp_device(
            self.h, d_dist.ptr if d_dist is not None else None, d_parent.ptr if d_parent is not None else None))

    def dyn_sssp(self, src, types, usrc, udst, batch_size):
        """The whole Dynamic function (spm_dyn_sssp): (dist, parent)."""
        usrc, udst = self._pair(usrc, udst)
        t = self._types(types, len(usrc))
        dist = np.empty(self.V, np.int32)
        parent = np.empty(self.V, np.int32)
        N.check("spm_dyn_sssp", N.device_lib().spm_dyn_sssp(
            self.h, int(src), t, N.ptr(usrc, C.c_int32), N.ptr(udst, C.c_int32), len(usrc), int(batch_size),
            N.ptr(dist, C.c_int32), N.ptr(parent, C.c_int32)))
        return dist, parent

    def dyntc_init(self):
        N.check("spm_dyntc_init", N.device_lib().spm_dyntc_init(self.h))

    def dyntc_batch(self, types, src, dst):
        src, dst = self._pair(src, dst)
        t = self._types(types, len(src))
        N.check("spm_dyntc_batch", N.device_lib().spm_dyntc_batch(
            self.h, t, N.ptr(src, C.c_int32), N.ptr(dst, C.c_int32), len(src)))

    def dyntc_count(self):
        c = C.c_int64(0)
        N.check("spm_dyntc_count", N.device_lib().spm_dyntc_count(self.h, C.byref(c)))
        return c.value

    def dyn_tc(self, types, usrc, udst, batch_size):
        """The whole Dynamic function (spm_dyn_tc): the final triangle count."""
        usrc, udst = self._pair(usrc, udst)
        t = self._types(types, len(usrc))
        c = C.c_int64(0)
        N.check("spm_dyn_tc", N.device_lib().spm_dyn_tc(
            self.h, t, N.ptr(usrc, C.c_int32), N.ptr(udst, C.c_int32), len(usrc), int(batch_size), C.byref(c)))
        return c.value

    def dynpr_init(self, beta, delta, maxIter):
        N.check("spm_dynpr_init", N.device_lib().spm_dynpr_init(self.h, beta, delta, int(maxIter)))

    def dynpr_batch(self, types, src, dst):
        src, dst = self._pair(src, dst)
        t = self._types(types, len(src))
        N.check("spm_dynpr_batch", N.device_lib().spm_dynpr_batch(
            self.h, t, N.ptr(src, C.c_int32), N.ptr(dst, C.c_int32), len(src)))

    def dynpr_result(self):
        pr = np.empty(self.V, np.float32)
        N.check("spm_dynpr_result", N.device_lib().spm_dynpr_result(self.h, N.ptr(pr, C.c_float)))
        return pr

    def dynpr_iterations(self):
        n = C.c_int32(0)
        N.check("spm_dynpr_iterations", N.device_lib().spm_dynpr_iterations(self.h, None, 0, C.byref(n)))
        out = np.zeros(max(n.value, 1), np.int32)
        N.check("spm_dynpr_iterations", N.device_lib().spm_dynpr_iterations(
            self.h, N.ptr(out, C.c_int32), n.value, C.byref(n)))
        return out[:n.value]

    def dyn_pr(self, beta, delta, maxIter, types, usrc, udst, batch_size):
        """The whole Dynamic function (spm_dyn_pr): the final ranks."""
        usrc, udst = self._pair(usrc, udst)
        t = self._types(types, len(usrc))
        pr = np.empty(self.V, np.float32)
        N.check("spm_dyn_pr", N.device_lib().spm_dyn_pr(
            self.h, beta, delta, int(maxIter), N.ptr(pr, C.c_float), t, N.ptr(usrc, C.c_int32),
            N.ptr(udst, C.c_int32), len(usrc), int(batch_size)))
        return pr

    # ---- schedules ----------------------------------------------------------
    def schedule(self):
        s = N.SpmSchedule()
        N.check("spm_get_schedule", N.device_lib().spm_get_schedule(self.h, C.byref(s)))
        return s

    def default_schedule(self):
        s = N.SpmSchedule()
        N.check("spm_default_schedule", N.device_lib().spm_default_schedule(self.h, C.byref(s)))
        return s

    def set_schedule(self, s=None, **kw):
        if s is None:
            s = self.schedule()
        for k, v in kw.items():
            setattr(s, k, int(v))
        N.check("spm_set_schedule", N.device_lib().spm_set_schedule(self.h, C.byref(s)))
        return s

    # ---- profiling -----------------------------------------------------------
    def profile(self, enable=True):
        N.check("spm_profile_enable", N.device_lib().spm_profile_enable(self.h, int(enable)))

    def profile_reset(self):
        N.check("spm_profile_reset", N.device_lib().spm_profile_reset(self.h))

    def profile_query(self, kernel):
        ms = C.c_double(0)
        n = C.c_int64(0)
        N.check("spm_profile_query", N.device_lib().spm_profile_query(self.h, kernel.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    # ---- algorithms, host buffers -----------------------------------------------
    def pr(self, beta, delta, maxIter, out=None):
        out = np.empty(self.V, np.float32) if out is None else out
        it = C.c_int32(0)
        N.check("spm_compute_pr", N.device_lib().spm_compute_pr(
            self.h, beta, delta, maxIter, N.ptr(out, C.c_float), C.byref(it)))
        return out, it.value

    def sssp(self, src, weight=None, out=None):
        out = np.empty(self.V, np.int32) if out is None else out
        w = None if weight is None else np.ascontiguousarray(weight, np.int32)
        N.check("spm_compute_sssp", N.device_lib().spm_compute_sssp(
            self.h, N.ptr(out, C.c_int32), N.ptr(w, C.c_int32), src))
        return out

    def tc(self):
        c = C.c_uint64(0)
        N.check("spm_compute_tc", N.device_lib().spm_compute_tc(self.h, C.byref(c)))
        return c.value

    def bc(self, sources, out=None):
        out = np.empty(self.V, np.float32) if out is None else out
        s = np.ascontiguousarray(sources, np.int32)
        N.check("spm_compute_bc", N.device_lib().spm_compute_bc(
            self.h, N.ptr(out, C.c_float), N.ptr(s, C.c_int32), len(s)))
        return out

    def bc_paths(self, src):
        lev = np.empty(self.V, np.int32)
        sig = np.empty(self.V, np.float64)
        N.check("spm_bc_paths", N.device_lib().spm_bc_paths(
            self.h, src, N.ptr(lev, C.c_int32), N.ptr(sig, C.c_double)))
        return lev, sig

    # ---- algorithms, device buffers ------------------------------------------------
    def pr_device(self, beta, delta, maxIter, d_out):
        it = C.c_int32(0)
        N.check("spm_pr_device", N.device_lib().spm_pr_device(self.h, beta, delta, maxIter, d_out.ptr, C.byref(it)))
        return it.value

    def pr_multi_device(self, bases, beta, delta, maxIter, d_out):
        """Batched PageRank jobs on a disjoint union (job k = vertices
        [bases[k], bases[k+1])); returns each job's iteration count."""
        b = np.ascontiguousarray(bases, np.int32)
        n = len(b) - 1
        it = np.zeros(max(n, 1), np.int32)
        N.check("spm_pr_multi_device", N.device_lib().spm_pr_multi_device(
            self.h, N.ptr(b, C.c_int32), n, beta, delta, maxIter, d_out.ptr, N.ptr(it, C.c_int32)))
        return it[:n]

    def sssp_device(self, src, d_out):
        r = C.c_int32(0)
        N.check("spm_sssp_device", N.device_lib().spm_sssp_device(self.h, src, d_out.ptr, C.byref(r)))
        return r.value

    def sssp_multi_device(self, sources, d_out):
        s = np.ascontiguousarray(sources, np.int32)
        r = C.c_int32(0)
        N.check("spm_sssp_multi_device", N.device_lib().spm_sssp_multi_device(
            self.h, N.ptr(s, C.c_int32), len(s), d_out.ptr, C.byref(r)))
        return r.value

    def bc_device(self, sources, d_out):
        s = np.ascontiguousarray(sources, np.int32)
        N.check("spm_bc_device", N.device_lib().spm_bc_device(self.h, N.ptr(s, C.c_int32), len(s), d_out.ptr))
