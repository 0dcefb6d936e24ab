"""spm_pr_device_noalloc: PageRank out of the library's file-scope device
array instead of the handle's workspace. Called through the C ABI and checked
against spm_pr_device on the same handle -- ranks and iteration count
bit-identical (same bins, same kernels, same summation order) -- on two calls
in a row, the second on a smaller graph, so nothing depends on what the first
call left in the shared array."""
import ctypes as C

import numpy as np
import pytest

from starplat_amd import _native as N
from starplat_amd import graph, hip

pytestmark = pytest.mark.gpu


def noalloc(dev, beta, delta, max_iter, d_out):
    it = C.c_int32(0)
    rc = N.device_lib().spm_pr_device_noalloc(dev.h, beta, delta, max_iter, d_out.ptr, C.byref(it))
    return rc, it.value


def test_noalloc_matches_pr_device_on_two_calls():
    big = graph.rmat(16, 16, seed=1)     # the PageRank configuration's graph (65,529 vertices)
    small = graph.rmat(11, 8, seed=3)    # then a smaller one
    for g in (big, small):
        dev = g.device()
        ref = hip.DeviceArray(g.V, np.float32)
        got = hip.DeviceArray(g.V, np.float32)
        it_ref = dev.pr_device(1e-10, 0.85, 100, ref)
        held = dev.device_bytes()
        rc, it = noalloc(dev, 1e-10, 0.85, 100, got)
        assert rc == 0, N.device_lib().spm_last_error()
        assert dev.device_bytes() == held  # the handle's workspace did not grow
        print(f"V={g.V}: {it} iterations (spm_pr_device {it_ref})")
        assert it == it_ref
        np.testing.assert_array_equal(got.to_host(), ref.to_host())
        # fewer iterations than convergence needs: the maxIter exit
        rc, it = noalloc(dev, 1e-10, 0.85, 7, got)
        assert rc == 0 and it == 7
        dev.pr_device(1e-10, 0.85, 7, ref)
        np.testing.assert_array_equal(got.to_host(), ref.to_host())
        ref.free()
        got.free()
        g.release_device()


def test_noalloc_refuses_what_needs_memory():
    # more vertices than the array holds (2^20 + 1: one edge reaching id 2^20)
    g = graph.from_edges(np.array([0], np.int32), np.array([1 << 20], np.int32))
    out = hip.DeviceArray(g.V, np.float32)
    rc, _ = noalloc(g.device(), 1e-10, 0.85, 10, out)
    assert rc == -1 and b"2^20" in N.device_lib().spm_last_error()
    out.free()
    g.release_device()
    # pending edge updates would need a fold (which allocates)
    g = graph.rmat(10, 8, seed=2)
    dev = g.device()
    src0 = np.repeat(np.arange(g.V, dtype=np.int32), np.diff(g.indexofNodes))[:1]
    dev.delete_edges(src0, g.edgeList[:1])
    out = hip.DeviceArray(g.V, np.float32)
    rc, _ = noalloc(dev, 1e-10, 0.85, 10, out)
    assert rc == -1 and b"pending" in N.device_lib().spm_last_error()
    dev.compact()
    rc, _ = noalloc(dev, 1e-10, 0.85, 10, out)
    assert rc == 0
    out.free()
    g.release_device()
