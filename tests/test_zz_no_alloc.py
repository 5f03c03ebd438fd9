"""The allocation-free block inverse (tfi_block_inverse_noalloc) against the
allocating entry point tfi_block_inverse on the same diagonal blocks: two
calls in a row, the second on a smaller mesh, both through the C-ABI.  The
device buffers are allocated here, by the test, with the HIP runtime; the
entry point itself takes device pointers and allocates nothing."""
import ctypes
import os

import numpy as np
import pytest

from tubularflowiga_amd import casegen

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    for path in ("/opt/rocm/lib/libamdhip64.so", "libamdhip64.so"):
        try:
            lib = ctypes.CDLL(path)
        except OSError:
            continue
        lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        lib.hipFree.argtypes = [ctypes.c_void_p]
        lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        lib.hipDeviceSynchronize.argtypes = []
        return lib
    pytest.fail("HIP runtime library not found")


class DeviceArray:
    def __init__(self, hip, host: np.ndarray):
        self.hip, self.nbytes = hip, host.nbytes
        self.ptr = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), self.nbytes) == 0
        assert hip.hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, H2D) == 0

    def download(self, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H) == 0
        return out

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


def _diagonal_values(dmesh, seed):
    """Random blocks with dominant diagonal blocks; a few diagonal blocks are
    made singular (all zero, or a rank-deficient block with a non-zero
    diagonal) so that the fallback and its count are exercised."""
    rng = np.random.default_rng(seed)
    _, _, diag = dmesh.pattern()
    vals = rng.standard_normal((dmesh.nnzb, 4, 4))
    vals[diag] += 8.0 * np.eye(4)
    zero = rng.choice(dmesh.nodes, size=3, replace=False)
    vals[diag[zero]] = 0.0
    rank = np.setdiff1d(rng.choice(dmesh.nodes, size=4, replace=False), zero)
    for n in rank:
        v = rng.standard_normal(4)
        vals[diag[n]] = np.outer(v, v) + np.diag(np.full(4, 1e-20))
    return vals.ravel()


@pytest.fixture(scope="module")
def ctx():
    from tubularflowiga_amd import device
    return device.Context(0)


def _noalloc(hip, dmesh, values):
    inv_host = np.full(16 * dmesh.nodes, np.nan)
    dv, di = DeviceArray(hip, values), DeviceArray(hip, inv_host)
    try:
        singular = ctypes.c_int(-1)
        dmesh.ctx.check(dmesh.lib.tfi_block_inverse_noalloc(dmesh.handle, dv.ptr, di.ptr, ctypes.byref(singular)))
        return di.download(inv_host).reshape(-1, 4, 4), singular.value
    finally:
        dv.free()
        di.free()


def test_noalloc_block_inverse_matches_the_allocating_entry_point(ctx):
    from tubularflowiga_amd import device
    hip = _hip()
    large = device.DeviceMesh(ctx, casegen.workload("ybif_small").mesh)
    small = device.DeviceMesh(ctx, casegen.workload("tube_c0").mesh)
    assert small.nodes < large.nodes
    try:
        for seed, dm in ((1, large), (2, small)):
            print(f"no-alloc block inverse: seed={seed} nodes={dm.nodes}", flush=True)
            values = _diagonal_values(dm, seed)
            got, got_singular = _noalloc(hip, dm, values)
            want, want_singular = dm.block_inverse(values)
            assert got_singular == want_singular and want_singular >= 3
            assert np.all(np.isfinite(got))
            assert np.max(np.abs(got - want)) <= 1e-14 * np.max(np.abs(want))
    finally:
        small.close()
        large.close()
