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
