"""Texture3D (texture_3d.h:6-10, texture_3d.cpp:3-51) through the C ABI.

create_tex_3d allocates the levels of a w x h x d texture; clear_tex_3d is
glClearTexImage of level 0 only (texture_3d.cpp:42), so higher levels keep
their contents; mip_tex_3d is glGenerateMipmap (texture_3d.cpp:49): each
level is the 2x2x2 box average of the level below, edges of odd size clamp
to the last texel.  The box filter is restated here in float32 numpy, summed
in the same order, and rounded to the storage type; the GPU must match it
bit for bit.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from voxel_cone_tracing_amd import _lib as L
from voxel_cone_tracing_amd.context import Context

pytestmark = pytest.mark.gpu


def _proto(lib):
    lib.vct_tex3d_create.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.POINTER(C.c_uint32)]
    lib.vct_tex3d_destroy.argtypes = [C.c_void_p, C.c_uint32]
    lib.vct_tex3d_activate.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32]
    lib.vct_tex3d_clear.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(C.c_float)]
    lib.vct_tex3d_mip.argtypes = [C.c_void_p, C.c_uint32]
    lib.vct_tex3d_upload.argtypes = [C.c_void_p, C.c_uint32, C.c_int, C.c_void_p, C.c_size_t]
    lib.vct_tex3d_download.argtypes = [C.c_void_p, C.c_uint32, C.c_int, C.c_void_p, C.c_size_t]


def level_dims(w, h, d, levels):
    out = []
    for _ in range(levels):
        out.append((w, h, d))
        w, h, d = max(1, w // 2), max(1, h // 2), max(1, d // 2)
    return out


def encode(f, dtype):
    """float32 (..., 4) -> the stored elements."""
    if dtype == L.FP32:
        return np.ascontiguousarray(f, np.float32)
    if dtype == L.FP16:
        return np.ascontiguousarray(f, np.float32).astype(np.float16).view(np.uint16)
    q = np.rint(np.clip(f, 0, 1).astype(np.float32) * np.float32(255.0)).astype(np.uint32)  # ties to even
    return (q[..., 0] | (q[..., 1] << 8) | (q[..., 2] << 16) | (q[..., 3] << 24)).astype(np.uint32)


def decode(a, dtype):
    if dtype == L.FP32:
        return a.astype(np.float32)
    if dtype == L.FP16:
        return a.view(np.float16).astype(np.float32)
    return np.stack([((a >> (8 * c)) & 255).astype(np.float32) / np.float32(255.0) for c in range(4)], -1)


def box_filter(src, dst_dims):
    """One glGenerateMipmap level: src (d, h, w, 4) float32 -> dst, the 8
    texels of the 2x2x2 block summed in x, y, z order, times 1/8."""
    sd, sh, sw = src.shape[:3]
    dw, dh, dd = dst_dims
    z, y, x = np.meshgrid(np.arange(dd), np.arange(dh), np.arange(dw), indexing="ij")
    acc = np.zeros((dd, dh, dw, 4), np.float32)
    for k in range(8):
        sx = np.minimum(2 * x + (k & 1), sw - 1)
        sy = np.minimum(2 * y + ((k >> 1) & 1), sh - 1)
        sz = np.minimum(2 * z + (k >> 2), sd - 1)
        acc = (acc + src[sz, sy, sx]).astype(np.float32)
    return (acc * np.float32(0.125)).astype(np.float32)


def elem_shape(dims, dtype):
    w, h, d = dims
    return (d, h, w) if dtype == L.RGBA8 else (d, h, w, 4)


def elem_dtype(dtype):
    return {L.RGBA8: np.uint32, L.FP16: np.uint16, L.FP32: np.float32}[dtype]


@pytest.mark.parametrize("dtype", [L.RGBA8, L.FP16, L.FP32])
@pytest.mark.parametrize("dims,levels", [((16, 8, 4), 5), ((7, 5, 3), 3)])
def test_texture3d_clear_and_mip(gpu_available, dtype, dims, levels):
    print(f"texture3d dtype={dtype} dims={dims} levels={levels}", flush=True)
    rng = np.random.default_rng(dtype * 10 + levels)
    ld = level_dims(*dims, levels)
    with Context(16, 16) as ctx:
        lib = ctx.lib
        _proto(lib)
        tex = C.c_uint32()
        L.check(lib.vct_tex3d_create(ctx.h, dims[0], dims[1], dims[2], levels, dtype, C.byref(tex)))
        assert tex.value >= 1  # GL names start at 1
        L.check(lib.vct_tex3d_activate(ctx.h, tex, 2))

        def up(level, a):
            a = np.ascontiguousarray(a)
            L.check(lib.vct_tex3d_upload(ctx.h, tex, level, a.ctypes.data_as(C.c_void_p), a.nbytes))

        def down(level):
            a = np.zeros(elem_shape(ld[level], dtype), elem_dtype(dtype))
            L.check(lib.vct_tex3d_download(ctx.h, tex, level, a.ctypes.data_as(C.c_void_p), a.nbytes))
            return a

        # every level starts at zero
        for l in range(levels):
            assert not down(l).any()
        # upload random levels, clear: level 0 only takes the colour
        lv = [encode(rng.random(elem_shape(ld[l], L.FP32)).astype(np.float32), dtype) for l in range(levels)]
        for l in range(levels):
            up(l, lv[l])
        colour = (C.c_float * 4)(0.25, 0.5, 0.75, 1.0)
        L.check(lib.vct_tex3d_clear(ctx.h, tex, colour))
        want0 = encode(np.broadcast_to(np.array([0.25, 0.5, 0.75, 1.0], np.float32),
                                       elem_shape(ld[0], L.FP32)), dtype)
        assert np.array_equal(down(0), want0)
        for l in range(1, levels):
            assert np.array_equal(down(l), lv[l]), f"clear touched level {l}"
        # the mip chain from a new level 0
        up(0, lv[0])
        L.check(lib.vct_tex3d_mip(ctx.h, tex))
        src = decode(lv[0], dtype)
        for l in range(1, levels):
            want = encode(box_filter(src, ld[l]), dtype)
            got = down(l)
            assert np.array_equal(got, want), f"level {l} differs"
            src = decode(got, dtype)
        L.check(lib.vct_tex3d_destroy(ctx.h, tex))
        L.check(lib.vct_tex3d_destroy(ctx.h, tex))  # unknown names are ignored, like glDeleteTextures
