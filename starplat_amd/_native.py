"""ctypes bindings of the package's two native libraries.

libstarplat_amd.so  gfx950 HIP kernels behind the C ABI of include/starplat_amd.h
libspm_host.so      host graph data layer of include/starplat_host.h

Both are built in-tree by ``starplat_amd.build.build()``. Loading fails loudly
when a library is missing: there is no CPU fallback for the device library.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DEVICE = os.path.join(_HERE, "libstarplat_amd.so")
LIB_HOST = os.path.join(_HERE, "libspm_host.so")

i32 = C.c_int32
i64 = C.c_int64
u64 = C.c_uint64
f32 = C.c_float
f64 = C.c_double
vp = C.c_void_p
P_i32 = C.POINTER(C.c_int32)
P_u64 = C.POINTER(C.c_uint64)
P_f32 = C.POINTER(C.c_float)
P_f64 = C.POINTER(C.c_double)
P_i64 = C.POINTER(C.c_int64)

SPM_OK = 0


class SpmSchedule(C.Structure):
    """Mirror of ``spm_schedule`` (include/starplat_amd.h)."""

    _fields_ = [
        ("pr_t1", i32),
        ("pr_t2", i32),
        ("pr_t3", i32),
        ("pr_group", i32),
        ("sssp_mode", i32),
        ("sssp_rounds", i32),
        ("sssp_grid", i32),
        ("tc_mode", i32),
        ("pr_mode", i32),
        ("pr_seg_log2", i32),
        ("relabel", i32),
        ("sssp_band_rounds", i32),
        ("reserved", i32 * 4),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class SpmhGraph(C.Structure):
    """Mirror of ``spmh_graph`` (include/starplat_host.h)."""

    _fields_ = [
        ("V", i32),
        ("E", i64),
        ("indexofNodes", P_i32),
        ("edgeList", P_i32),
        ("edgeLen", P_i32),
        ("rev_indexofNodes", P_i32),
        ("srcList", P_i32),
    ]


P_spmh = C.POINTER(SpmhGraph)

# (name, restype, argtypes) of every function in include/starplat_amd.h
DEVICE_API = [
    ("spm_last_error", C.c_char_p, []),
    ("spm_version", C.c_int, []),
    ("spm_graph_create", C.c_int, [i32, i64, P_i32, P_i32, P_i32, P_i32, P_i32, C.POINTER(vp)]),
    ("spm_graph_destroy", C.c_int, [vp]),
    ("spm_graph_set_weights", C.c_int, [vp, P_i32]),
    ("spm_graph_stream", vp, [vp]),
    ("spm_graph_sync", C.c_int, [vp]),
    ("spm_graph_device_bytes", i64, [vp]),
    ("spm_compute_pr", C.c_int, [vp, f32, f32, i32, P_f32, P_i32]),
    ("spm_pr_device", C.c_int, [vp, f32, f32, i32, vp, P_i32]),
    ("spm_pr_device_noalloc", C.c_int, [vp, f32, f32, i32, vp, P_i32]),
    ("spm_pr_multi_device", C.c_int, [vp, P_i32, i32, f32, f32, i32, vp, P_i32]),
    ("spm_compute_sssp", C.c_int, [vp, P_i32, P_i32, i32]),
    ("spm_sssp_device", C.c_int, [vp, i32, vp, P_i32]),
    ("spm_sssp_multi_device", C.c_int, [vp, P_i32, i32, vp, P_i32]),
    ("spm_compute_tc", C.c_int, [vp, P_u64]),
    ("spm_compute_bc", C.c_int, [vp, P_f32, P_i32, i32]),
    ("spm_bc_device", C.c_int, [vp, P_i32, i32, vp]),
    ("spm_bc_paths", C.c_int, [vp, i32, P_i32, P_f64]),
    ("spm_graph_delete_edges", C.c_int, [vp, P_i32, P_i32, i32]),
    ("spm_graph_add_edges", C.c_int, [vp, P_i32, P_i32, P_i32, i32]),
    ("spm_graph_compact", C.c_int, [vp]),
    ("spm_graph_num_edges", C.c_int, [vp, P_i64]),
    ("spm_graph_export", C.c_int, [vp, P_i32, P_i32, P_i32, P_i32, P_i32]),
    ("spm_dynsssp_init", C.c_int, [vp, i32]),
    ("spm_dynsssp_batch", C.c_int, [vp, C.c_char_p, P_i32, P_i32, i32]),
    ("spm_dynsssp_result", C.c_int, [vp, P_i32, P_i32]),
    ("spm_dynsssp_device", C.c_int, [vp, vp, vp]),
    ("spm_dyn_sssp", C.c_int, [vp, i32, C.c_char_p, P_i32, P_i32, i32, i32, P_i32, P_i32]),
    ("spm_dyntc_init", C.c_int, [vp]),
    ("spm_dyntc_batch", C.c_int, [vp, C.c_char_p, P_i32, P_i32, i32]),
    ("spm_dyntc_count", C.c_int, [vp, P_i64]),
    ("spm_dyn_tc", C.c_int, [vp, C.c_char_p, P_i32, P_i32, i32, i32, P_i64]),
    ("spm_dynpr_init", C.c_int, [vp, f32, f32, i32]),
    ("spm_dynpr_batch", C.c_int, [vp, C.c_char_p, P_i32, P_i32, i32]),
    ("spm_dynpr_result", C.c_int, [vp, P_f32]),
    ("spm_dynpr_iterations", C.c_int, [vp, P_i32, i32, P_i32]),
    ("spm_dyn_pr", C.c_int, [vp, f32, f32, i32, P_f32, C.c_char_p, P_i32, P_i32, i32, i32]),
    ("spm_get_schedule", C.c_int, [vp, C.POINTER(SpmSchedule)]),
    ("spm_set_schedule", C.c_int, [vp, C.POINTER(SpmSchedule)]),
    ("spm_default_schedule", C.c_int, [vp, C.POINTER(SpmSchedule)]),
    ("spm_mlp_create", C.c_int, [i32, i32, i32, u64, C.POINTER(vp)]),
    ("spm_mlp_destroy", C.c_int, [vp]),
    ("spm_mlp_num_params", C.c_int, []),
    ("spm_mlp_set_params", C.c_int, [vp, P_f32]),
    ("spm_mlp_get_params", C.c_int, [vp, P_f32]),
    ("spm_mlp_forward", C.c_int, [vp, P_f32, i32, P_f32]),
    ("spm_mlp_predict", C.c_int, [vp, P_f32, i32, P_i32]),
    ("spm_mlp_grad", C.c_int, [vp, P_f32, P_i32, i32, P_f32, P_f32]),
    ("spm_mlp_train", C.c_int, [vp, P_f32, P_i32, i32, i32, i32, f32, u64, P_f32]),
    ("spm_profile_enable", C.c_int, [vp, C.c_int]),
    ("spm_profile_reset", C.c_int, [vp]),
    ("spm_profile_query", C.c_int, [vp, C.c_char_p, P_f64, P_i64]),
    ("spm_gather_rate", C.c_int, [i32, i32, P_f64]),
]

HOST_API = [
    ("spmh_rmat", P_spmh, [C.c_int, C.c_int, f64, f64, f64, u64, C.c_int, C.c_int, C.c_int]),
    ("spmh_er", P_spmh, [i64, i64, u64, C.c_int, C.c_int, C.c_int]),
    ("spmh_grid", P_spmh, [C.c_int, C.c_int, u64, C.c_int, C.c_int]),
    ("spmh_from_edges", P_spmh, [i64, P_i32, P_i32, P_i32]),
    ("spmh_parse_text", P_spmh, [C.c_char_p]),
    ("spmh_union", P_spmh, [C.POINTER(P_spmh), i32, P_i64]),
    ("spmh_write_text", C.c_int, [P_spmh, C.c_char_p, C.c_int]),
    ("spmh_write_binary", C.c_int, [P_spmh, C.c_char_p]),
    ("spmh_read_binary", P_spmh, [C.c_char_p]),
    ("spmh_free", None, [P_spmh]),
    ("spmh_sources", C.c_int, [i32, i32, u64, P_i32]),
    ("spmh_edge_map", C.c_int, [i32, P_i32, P_i32, P_i32, P_i32, P_i32]),
    ("spmh_features", C.c_int, [P_spmh, f32, P_f32]),
    ("spmh_set_threads", None, [C.c_int]),
    ("spmh_model_pr_bytes", f64, [P_spmh]),
    ("spmh_model_sssp", f64, [P_spmh, P_i32, i32, P_i64, P_i64, P_i64]),
    ("spmh_model_tc_bytes", f64, [P_spmh, P_i64, P_i64]),
    ("spmh_model_bc_bytes", f64, [P_spmh, P_i32, i32, P_i64]),
    ("spmh_bc_gather_profile", C.c_int, [P_spmh, P_i32, i32, P_f64]),
]

_dev = None
_host = None


def _bind(lib, table):
    for name, res, args in table:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def device_lib():
    """The HIP kernel library; raises if it has not been built."""
    global _dev
    if _dev is None:
        if not os.path.exists(LIB_DEVICE):
            raise RuntimeError(
                f"{LIB_DEVICE} is missing: build it with "
                "`python -c 'import __graft_entry__ as m; m.build()'` (there is no CPU fallback)")
        _dev = _bind(C.CDLL(LIB_DEVICE, mode=C.RTLD_GLOBAL), DEVICE_API)
    return _dev


def host_lib():
    global _host
    if _host is None:
        if not os.path.exists(LIB_HOST):
            raise RuntimeError(f"{LIB_HOST} is missing: build the package first")
        _host = _bind(C.CDLL(LIB_HOST), HOST_API)
    return _host


class SpmError(RuntimeError):
    def __init__(self, fn, code):
        msg = device_lib().spm_last_error().decode(errors="replace")
        super().__init__(f"{fn} failed with status {code}: {msg}")
        self.code = code


def check(fn, code):
    if code != SPM_OK:
        raise SpmError(fn, code)


def ptr(a, ctype):
    """ctypes pointer to a contiguous numpy array (or None)."""
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))
