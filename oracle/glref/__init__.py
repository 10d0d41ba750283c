"""Python side of the llvmpipe reference harness (test infrastructure only).

Runs the reference's own shader files (read from REFERENCE_SHADERS at run
time) under Mesa llvmpipe through oracle/_ref/libvct_glref.so, built from
oracle/glref/glref.cpp by oracle/glref/Makefile.  Available only where both
the reference tree and Mesa's swrast driver exist (this container).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE.parent / "_ref" / "libvct_glref.so"
REFERENCE_SHADERS = Path(os.environ.get("VCT_REFERENCE_SHADERS", "/root/reference/shader"))

_lib = None
_FP = C.POINTER(C.c_float)
_UP = C.POINTER(C.c_uint32)
_DP = C.POINTER(C.c_double)


def available() -> bool:
    return REFERENCE_SHADERS.is_dir() and Path("/usr/lib/x86_64-linux-gnu/dri/swrast_dri.so").exists()


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not available():
        raise RuntimeError("reference shaders or Mesa swrast driver not present")
    if not LIB.exists():
        subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    lib = C.CDLL(str(LIB))
    lib.glref_error.restype = C.c_char_p
    lib.glref_renderer.restype = C.c_char_p
    lib.glref_init.argtypes = [C.c_char_p]
    lib.glref_frame.argtypes = [C.c_int, C.c_float, _FP, _FP, _FP, C.c_int, C.c_int, C.c_int, _FP, _FP, _UP,
                                C.c_void_p, C.c_int, _FP, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                C.c_float, _UP, _UP, _FP, _FP, _FP, _DP]
    if lib.glref_init(str(REFERENCE_SHADERS).encode()) != 0:
        raise RuntimeError("llvmpipe harness: " + lib.glref_error().decode())
    _lib = lib
    return lib


def renderer() -> str:
    return load().glref_renderer().decode()


def _c(a, dt):
    return np.ascontiguousarray(a, dt)


def frame(res, cube_size, model, view, proj, W, H, pos, nrm, tri_mat, mats, lights,
          phases=(False, True, True, False), view_voxel_dir=7, view_voxel_lod=0.0,
          want=("level0", "mips", "image", "gbuffer")):
    """One reference frame.  lights: list of (position, colour, intensity, kind) point lights.
    Returns dict with level0 (res,res,res) u32, mips {(l,d): (s,s,s) u32}, image (H,W,4),
    gpos/gnrm (H,W,4), times [voxelize, mip, trace] seconds."""
    lib = load()
    model = _c(model, np.float32).reshape(16)
    view = _c(view, np.float32).reshape(16)
    proj = _c(proj, np.float32).reshape(16)
    pos = _c(pos, np.float32)
    nrm = _c(nrm, np.float32)
    ntri = pos.reshape(-1, 9).shape[0]
    tm = None if tri_mat is None else _c(tri_mat, np.uint32)
    mb = _c(mats, np.float32).reshape(-1, 32)
    lb = np.zeros((max(1, len(lights)), 8), np.float32)
    for i, (p, c, it, k) in enumerate(lights):
        assert k == 0, "the reference shaders have point lights only"
        lb[i, 0:3], lb[i, 3:6], lb[i, 6] = p, c, it
    out = {}
    l0 = np.zeros(res ** 3, np.uint32) if "level0" in want else None
    nm = sum(6 * (res >> l) ** 3 for l in range(1, 7) if res >> l)
    mp = np.zeros(nm, np.uint32) if "mips" in want else None
    img = np.zeros((H, W, 4), np.float32) if "image" in want else None
    gp = np.zeros((H, W, 4), np.float32) if "gbuffer" in want else None
    gn = np.zeros((H, W, 4), np.float32) if "gbuffer" in want else None
    times = np.zeros(3, np.float64)
    f = lambda a: None if a is None else a.ctypes.data_as(_FP)  # noqa: E731
    u = lambda a: None if a is None else a.ctypes.data_as(_UP)  # noqa: E731
    rc = lib.glref_frame(res, cube_size, f(model), f(view), f(proj), W, H, ntri, f(pos), f(nrm), u(tm),
                         mb.ctypes.data, mb.shape[0], f(lb), len(lights), *(int(x) for x in phases),
                         view_voxel_dir, view_voxel_lod, u(l0), u(mp), f(img), f(gp), f(gn),
                         times.ctypes.data_as(_DP))
    if rc != 0:
        raise RuntimeError("llvmpipe frame failed: " + lib.glref_error().decode())
    if l0 is not None:
        out["level0"] = l0.reshape(res, res, res)
    if mp is not None:
        mips, off = {}, 0
        for l in range(1, 7):
            s = res >> l
            if s == 0:
                break
            for d in range(6):
                mips[(l, d)] = mp[off:off + s ** 3].reshape(s, s, s)
                off += s ** 3
        out["mips"] = mips
    if img is not None:
        out["image"] = img
    if gp is not None:
        out["gpos"], out["gnrm"] = gp, gn
    out["times"] = times
    return out
