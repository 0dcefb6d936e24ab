"""The C-ABI libraries load and export every function their headers declare
(no GPU calls)."""
import ctypes
import os
import re

from conftest import ROOT
from starplat_amd import _native as N


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(spmh?_[a-z0-9_]+)\s*\(", text)))


def test_device_library_exports_header():
    names = declared("starplat_amd.h")
    assert "spm_compute_sssp" in names and len(names) >= 20
    lib = ctypes.CDLL(N.LIB_DEVICE)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding table covers the header exactly
    assert sorted(n for n, _, _ in N.DEVICE_API) == names


def test_host_library_exports_header():
    names = declared("starplat_host.h")
    lib = ctypes.CDLL(N.LIB_HOST)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(n for n, _, _ in N.HOST_API) == names


def test_device_library_is_gfx950():
    data = open(N.LIB_DEVICE, "rb").read()
    assert b"gfx950" in data


def test_version_and_error_string_without_gpu():
    lib = N.device_lib()
    assert lib.spm_version() >= 100
    assert isinstance(lib.spm_last_error(), bytes)


def test_graph_create_rejects_malformed_csr_without_gpu():
    """A CSR whose offsets decrease, whose neighbour ids leave [0, V) or whose
    rows are not ascending is refused with SPM_EINVAL before any device call
    (kernels index by these values unchecked, and edge updates and the
    triangle count binary-search the rows)."""
    import numpy as np

    lib = N.device_lib()
    P = ctypes.POINTER(ctypes.c_int32)

    def create(off, adj, roff=None, radj=None):
        off = np.asarray(off, np.int32)
        adj = np.asarray(adj, np.int32)
        h = ctypes.c_void_p()
        args = [N.ptr(off, ctypes.c_int32), N.ptr(adj, ctypes.c_int32), None]
        if roff is not None:
            roff = np.asarray(roff, np.int32)
            radj = np.asarray(radj, np.int32)
            args += [N.ptr(roff, ctypes.c_int32), N.ptr(radj, ctypes.c_int32)]
        else:
            args += [ctypes.cast(None, P), ctypes.cast(None, P)]
        return lib.spm_graph_create(len(off) - 1, len(adj), *args, ctypes.byref(h)), h

    cases = [
        ([0, 2, 1, 3], [1, 2, 0]),            # offsets decrease
        ([0, 1, 2, 3], [1, 3, 0]),            # neighbour id == V
        ([0, 1, 2, 3], [1, -1, 0]),           # negative neighbour id
        ([0, 2, 3, 3], [2, 1, 0]),            # row 0 descending
        ([0, 3, 3, 3], [0, 2, 1]),            # row 0 out of order in its middle
        ([0, 2, 2, 2], [3, 1]),               # out of range and out of order
    ]
    for off, adj in cases:
        rc, h = create(off, adj)
        assert rc == -1 and not h.value, (off, adj)
        assert b"edgeList" in lib.spm_last_error()
    assert b"ascending" in (create([0, 2, 3, 3], [2, 1, 0])[0], lib.spm_last_error())[1]
    # a reverse CSR with an unsorted row
    rc, h = create([0, 1, 2, 3], [1, 2, 0], [0, 0, 3, 3], [2, 0, 1])
    assert rc == -1 and not h.value and b"srcList" in lib.spm_last_error()
    # a valid forward CSR with a bad reverse CSR
    rc, h = create([0, 1, 2, 3], [1, 2, 0], [0, 1, 2, 3], [2, 0, 7])
    assert rc == -1 and not h.value and b"srcList" in lib.spm_last_error()
