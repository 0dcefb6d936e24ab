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
        if weight.size != self.E:
            raise ValueError("weight must have one entry per CSR edge")
        N.check("spm_graph_set_weights", N.device_lib().spm_graph_set_weights(self.h, N.ptr(weight, C.c_int32)))

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
