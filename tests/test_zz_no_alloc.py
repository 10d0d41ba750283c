"""The entry point that allocates nothing (vct_mlp_forward_noalloc) against
vct_mlp_forward, through the C ABI.

Inputs and outputs are device buffers the test allocates itself (hipMalloc
through ctypes); the entry point stages the BF16 rows in the library's fixed
per-GPU array.  Two calls in a row, the second with fewer rows, must each
equal vct_mlp_forward's output bit for bit (the second call must not see
rows the first left behind); more rows than the array holds are rejected.
"""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

from voxel_cone_tracing_amd import _lib as L
from voxel_cone_tracing_amd.context import Context

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class DeviceBuffer:
    """hipMalloc'd device memory (the caller's buffers)."""

    hip = None

    def __init__(self, nbytes):
        if DeviceBuffer.hip is None:
            h = C.CDLL("libamdhip64.so")
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipFree.argtypes = [C.c_void_p]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            h.hipDeviceSynchronize.argtypes = []
            DeviceBuffer.hip = h
        self.p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.p), C.c_size_t(nbytes)) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0  # H2D

    def get(self, a):
        assert self.hip.hipMemcpy(a.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(a.nbytes), 2) == 0  # D2H
        return a

    def free(self):
        self.hip.hipFree(self.p)


def test_noalloc_forward_equals_forward(gpu_available):
    rng = np.random.default_rng(5)
    p = (rng.standard_normal(L.MLP_PARAMS) * 0.2).astype(np.float32)
    sizes = [5000, 1000]  # the second call smaller than the first
    with Context(16, 16) as ctx:
        ctx.mlp_set_params(p)
        ctx.sync()
        dx = DeviceBuffer(max(sizes) * 34 * 4)
        dy = DeviceBuffer(max(sizes) * 3 * 4)
        try:
            for n in sizes:
                say(f"no-alloc forward, n={n}")
                x = rng.random((n, 34)).astype(np.float32) * 2 - 0.5
                want = ctx.mlp_forward(x)
                dx.put(x)
                dy.put(np.full((max(sizes), 3), np.nan, np.float32))
                L.check(ctx.lib.vct_mlp_forward_noalloc(ctx.h, dx.p, n, dy.p, None))
                ctx.sync()
                got = dy.get(np.zeros((max(sizes), 3), np.float32))
                assert np.array_equal(got[:n].view(np.uint32), want.view(np.uint32)), "differs from vct_mlp_forward"
                assert np.isnan(got[n:]).all(), "wrote past the n rows"
            rc = ctx.lib.vct_mlp_forward_noalloc(ctx.h, dx.p, L.MLP_NOALLOC_MAX_ROWS + 1, dy.p, None)
            assert rc == L.VCT_ERR_LIMIT
        finally:
            dx.free()
            dy.free()
