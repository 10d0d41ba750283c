"""CPU oracle of the voxel cone tracing hot path (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this package, and only as the checker: the product never calls it.  The
C++ restatement lives in oracle/vct_oracle.cpp (see its header for the
reference file:line of every function); `make -C oracle` builds it.

What pins each function (golden vectors, closed forms, brute force, the
reference shaders under llvmpipe) is enforced by tests/test_oracle_*.py.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

_lib = None
_FP = C.POINTER(C.c_float)
_DP = C.POINTER(C.c_double)
_UP = C.POINTER(C.c_uint32)
_U64P = C.POINTER(C.c_uint64)
_I64P = C.POINTER(C.c_int64)


class TraceParams(C.Structure):
    _fields_ = [
        ("W", C.c_int), ("H", C.c_int), ("cube_size", C.c_float), ("cube_res", C.c_int),
        ("cam", C.c_float * 3), ("en_direct", C.c_int), ("en_diffuse", C.c_int), ("en_specular", C.c_int),
        ("en_shadow", C.c_int), ("view_dir", C.c_int), ("view_lod", C.c_float), ("cone_set", C.c_int),
        ("demodulate", C.c_int),
    ]


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        lib = C.CDLL(str(LIB))
        lib.orc_voxelize.restype = C.c_int64
        lib.orc_voxelize.argtypes = [C.c_int, C.c_float, _FP, C.c_int, _FP, _FP, _UP, C.c_void_p, C.c_void_p,
                                     C.c_int, _UP, _FP, _DP, _UP]
        if hasattr(lib, "orc_voxelize_draws"):
            lib.orc_voxelize_draws.restype = C.c_int64
            lib.orc_voxelize_draws.argtypes = [C.c_int, C.c_float, C.c_int, _FP, C.c_void_p, _FP, _FP, _UP,
                                               C.c_void_p, C.c_void_p, C.c_int, _UP, _FP, _DP, _UP]
        if hasattr(lib, "orc_voxel_view_f32"):
            lib.orc_voxel_view_f32.argtypes = [C.c_int, C.c_int, _FP, _FP, C.c_int, _FP, C.c_void_p, _FP, C.c_int,
                                               C.c_int, C.c_void_p, C.c_float, C.c_int, C.c_float, _FP]
        lib.orc_rgba8_avg_step.restype = C.c_uint32
        lib.orc_rgba8_avg_step.argtypes = [C.c_uint32, _FP]
        lib.orc_set_round_mode.argtypes = [C.c_int]
        lib.orc_mip_f32.argtypes = [C.c_int, C.c_int, C.c_int, _FP, _FP]
        lib.orc_mip_f64.argtypes = [C.c_int, C.c_int, _DP, _DP]
        lib.orc_gbuffer_raster.argtypes = [C.c_int, C.c_int, _FP, _FP, _FP, C.c_int, _FP, C.c_uint32, _U64P]
        lib.orc_gbuffer_resolve.argtypes = [C.c_int, C.c_int, _FP, _FP, _FP, C.c_int, _FP, _FP, _UP, C.c_uint32,
                                            _U64P, _FP, _FP, _DP, _DP]
        lib.orc_trace_f32.argtypes = [C.POINTER(TraceParams), C.c_int, C.c_int, C.POINTER(_FP), _FP, _FP,
                                      C.c_void_p, C.c_void_p, C.c_int, _I64P, C.c_int64, _FP, _I64P]
        lib.orc_trace_f64.argtypes = [C.POINTER(TraceParams), C.c_int, C.c_int, C.POINTER(_DP), _DP, _DP,
                                      C.c_void_p, C.c_void_p, C.c_int, _I64P, C.c_int64, _DP, _I64P]
        if hasattr(lib, "orc_trace_counts_f32"):
            lib.orc_trace_counts_f32.argtypes = lib.orc_trace_f32.argtypes
        for n in ("orc_store_round",):
            getattr(lib, n).argtypes = [_FP, C.c_int64, C.c_int]
        lib.orc_encode_f16.argtypes = [_FP, C.POINTER(C.c_uint16), C.c_int64]
        lib.orc_encode_bf16.argtypes = [_FP, C.POINTER(C.c_uint16), C.c_int64]
        lib.orc_decode_f16.argtypes = [C.POINTER(C.c_uint16), _FP, C.c_int64]
        lib.orc_decode_bf16.argtypes = [C.POINTER(C.c_uint16), _FP, C.c_int64]
        lib.orc_encode_rgba8.argtypes = [_FP, _UP, C.c_int64]
        lib.orc_decode_rgba8.argtypes = [_UP, _FP, C.c_int64]
        _lib = lib
    return _lib


def _f(a):
    return a.ctypes.data_as(_FP)


def _d(a):
    return a.ctypes.data_as(_DP)


def _u(a):
    return a.ctypes.data_as(_UP)


def _c(a, dt):
    return np.ascontiguousarray(a, dt)


def material_bytes(mats) -> np.ndarray:
    """Materials (ctypes Material structures or a (n, 32) float32 record array) as 128-byte records."""
    if isinstance(mats, np.ndarray):
        return _c(mats, np.float32).reshape(-1, 32)
    out = np.zeros((len(mats), 32), np.float32)
    for i, m in enumerate(mats):
        out[i] = np.frombuffer(bytes(m), np.float32)
    return out


def light_bytes(lights) -> np.ndarray:
    """lights: list of (position, color, intensity, kind)."""
    out = np.zeros((max(1, len(lights)), 8), np.float32)
    for i, (p, c, it, k) in enumerate(lights):
        out[i, 0:3] = p
        out[i, 3:6] = c
        out[i, 6] = it
        out[i, 7] = np.array([k], np.int32).view(np.float32)[0]
    return out


def set_round_mode(away: bool) -> None:
    load().orc_set_round_mode(1 if away else 0)


def rgba8_avg_step(stored: int, color) -> int:
    c = _c(color, np.float32)
    return int(load().orc_rgba8_avg_step(stored, _f(c)))


def voxelize(res, cube_size, M, pos, nrm, tri_mat, mats, lights, want=("ref", "q32", "f64", "count")):
    """One draw (model matrix M) voxelized into level 0."""
    return voxelize_draws(res, cube_size, [(M, pos, nrm, tri_mat)], mats, lights, want)


def voxelize_draws(res, cube_size, draws, mats, lights, want=("ref", "q32", "f64", "count")):
    """Draws [(M, pos, nrm, tri_mat or None), ...] in submission order
    voxelized into level 0 (orc_voxelize_draws)."""
    lib = load()
    Ms = np.ascontiguousarray(np.stack([np.asarray(d[0], np.float32).reshape(16) for d in draws]))
    counts = np.array([np.asarray(d[1]).reshape(-1, 9).shape[0] for d in draws], np.int32)
    pos = _c(np.concatenate([np.asarray(d[1], np.float32).reshape(-1) for d in draws]), np.float32)
    nrm = _c(np.concatenate([np.asarray(d[2], np.float32).reshape(-1) for d in draws]), np.float32)
    tm = None
    if any(d[3] is not None for d in draws):
        tm = _c(np.concatenate([np.zeros(c, np.uint32) if d[3] is None else np.asarray(d[3], np.uint32).reshape(-1)
                                for d, c in zip(draws, counts)]), np.uint32)
    mb = material_bytes(mats)
    lb = light_bytes(lights)
    n = res ** 3
    out = {}
    ref = np.zeros(n, np.uint32) if "ref" in want else None
    q32 = np.zeros((n, 4), np.float32) if "q32" in want else None
    f64 = np.zeros((n, 4), np.float64) if "f64" in want else None
    cnt = np.zeros(n, np.uint32) if "count" in want else None
    nf = lib.orc_voxelize_draws(res, cube_size, len(draws), _f(Ms), counts.ctypes.data, _f(pos), _f(nrm),
                                None if tm is None else _u(tm), mb.ctypes.data, lb.ctypes.data, len(lights),
                                None if ref is None else _u(ref), None if q32 is None else _f(q32),
                                None if f64 is None else _d(f64), None if cnt is None else _u(cnt))
    s = (res, res, res)
    if ref is not None:
        out["ref"] = ref.reshape(s)
    if q32 is not None:
        out["q32"] = q32.reshape(s + (4,))
    if f64 is not None:
        out["f64"] = f64.reshape(s + (4,))
    if cnt is not None:
        out["count"] = cnt.reshape(s)
    out["fragments"] = int(nf)
    return out


def _level_layout(res, levels):
    offs = {}
    off = 0
    for l in range(1, levels):
        for d in range(6):
            s = res >> l
            offs[(l, d)] = (off, s)
            off += s ** 3 * 4
    return offs, off


def mip_f32(res, levels, dtype, level0):
    """Levels 1..L-1 of the 6 maps, rounded to dtype per level: dict (l, d) -> (s,s,s,4)."""
    lib = load()
    l0 = _c(level0, np.float32)
    offs, total = _level_layout(res, levels)
    out = np.zeros(max(total, 1), np.float32)
    lib.orc_mip_f32(res, levels, dtype, _f(l0), _f(out))
    return {k: out[o:o + s ** 3 * 4].reshape(s, s, s, 4) for k, (o, s) in offs.items()}


def mip_f64(res, levels, level0):
    lib = load()
    l0 = _c(level0, np.float64)
    offs, total = _level_layout(res, levels)
    out = np.zeros(max(total, 1), np.float64)
    lib.orc_mip_f64(res, levels, _d(l0), _d(out))
    return {k: out[o:o + s ** 3 * 4].reshape(s, s, s, 4) for k, (o, s) in offs.items()}


def gbuffer(W, H, view, proj, draws):
    """draws: list of (M, pos, nrm, tri_mat).  Returns pos32, nrm32, pos64, nrm64 (H, W, 4) and keys."""
    lib = load()
    view = _c(view, np.float32).reshape(16)
    proj = _c(proj, np.float32).reshape(16)
    keys = np.full(W * H, np.uint64(0xFFFFFFFFFFFFFFFF), np.uint64)
    prepared = []
    base = 0
    for (M, pos, nrm, tm) in draws:
        M = _c(M, np.float32).reshape(16)
        pos = _c(pos, np.float32)
        nrm = _c(nrm, np.float32)
        tm = None if tm is None else _c(tm, np.uint32)
        n = pos.reshape(-1, 9).shape[0]
        lib.orc_gbuffer_raster(W, H, _f(view), _f(proj), _f(M), n, _f(pos), base, keys.ctypes.data_as(_U64P))
        prepared.append((M, pos, nrm, tm, n, base))
        base += n
    p32 = np.zeros((H, W, 4), np.float32)
    n32 = np.zeros((H, W, 4), np.float32)
    p64 = np.zeros((H, W, 4), np.float64)
    n64 = np.zeros((H, W, 4), np.float64)
    for (M, pos, nrm, tm, n, b) in prepared:
        lib.orc_gbuffer_resolve(W, H, _f(view), _f(proj), _f(M), n, _f(pos), _f(nrm),
                                None if tm is None else _u(tm), b, keys.ctypes.data_as(_U64P),
                                _f(p32), _f(n32), _d(p64), _d(n64))
    return p32, n32, p64, n64, keys.reshape(H, W)


def _params(W, H, cube_size, res, cam, phases=(False, True, True, False), view_dir=7, view_lod=0.0, cone_set=0,
            demodulate=False):
    p = TraceParams()
    p.W, p.H, p.cube_size, p.cube_res = W, H, cube_size, res
    for k in range(3):
        p.cam[k] = cam[k]
    p.en_direct, p.en_diffuse, p.en_specular, p.en_shadow = (int(x) for x in phases)
    p.view_dir, p.view_lod, p.cone_set, p.demodulate = view_dir, view_lod, cone_set, int(demodulate)
    return p


def trace(dtype_float, res, levels, level0, mips, gpos, gnrm, mats, lights, cam, cube_size, cone_set=0,
          phases=(False, True, True, False), pixels=None, view_dir=7, view_lod=0.0, demodulate=False,
          early_exit_steps=False):
    """Shade pixels.  dtype_float: np.float32 (faithful) or np.float64 (ideal).
    level0: (res,res,res,4); mips: dict (l, d) -> (s,s,s,4); gpos/gnrm: (H, W, 4).
    Returns (out (npix, 4), steps): the cone steps of the reference loop, or
    (FP32 with early_exit_steps) the pair (reference steps, steps of the loop
    that also stops once every later sample reads the zero border)."""
    lib = load()
    f64 = dtype_float == np.float64
    dt = np.float64 if f64 else np.float32
    H, W = gpos.shape[:2]
    keep = [_c(level0, dt)]
    ptrs = []
    P = C.POINTER(C.c_double) if f64 else C.POINTER(C.c_float)
    for l in range(levels):
        for d in range(6):
            a = keep[0] if l == 0 else _c(mips[(l, d)], dt)
            keep.append(a)
            ptrs.append(a.ctypes.data_as(P))
    arr = (P * len(ptrs))(*ptrs)
    gp = _c(gpos, dt)
    gn = _c(gnrm, dt)
    mb = material_bytes(mats)
    lb = light_bytes(lights)
    if pixels is None:
        pix = None
        npix = W * H
    else:
        pix = _c(pixels, np.int64)
        npix = pix.shape[0]
    out = np.zeros((npix, 4), dt)
    steps = (C.c_int64 * 2)(0, 0)
    prm = _params(W, H, cube_size, res, cam, phases, view_dir, view_lod, cone_set, demodulate)
    if early_exit_steps and f64:
        raise ValueError("early_exit_steps is counted by the FP32 oracle")
    fn = lib.orc_trace_f64 if f64 else (lib.orc_trace_counts_f32 if early_exit_steps else lib.orc_trace_f32)
    fn(C.byref(prm), res, levels, arr, gp.ctypes.data_as(P), gn.ctypes.data_as(P), mb.ctypes.data, lb.ctypes.data,
       len(lights), None if pix is None else pix.ctypes.data_as(_I64P), npix, out.ctypes.data_as(P),
       C.cast(steps, _I64P))
    return out, ((steps[0], steps[1]) if early_exit_steps else steps[0])


def voxel_view(W, H, view, proj, draws, res, levels, level0, mips, cube_size, view_dir, view_lod):
    """The voxel view with GL's in-order SRC_ALPHA blending (orc_voxel_view_f32):
    draws [(M, pos), ...] in submission order; level0 (res,res,res,4) and mips
    dict (l, d) -> (s,s,s,4) decoded volume.  Returns the (H, W, 4) image."""
    lib = load()
    keep = [_c(level0, np.float32)]
    ptrs = []
    P = C.POINTER(C.c_float)
    for l in range(levels):
        for d in range(6):
            a = keep[0] if l == 0 else _c(mips[(l, d)], np.float32)
            keep.append(a)
            ptrs.append(a.ctypes.data_as(P))
    arr = (P * len(ptrs))(*ptrs)
    Ms = np.ascontiguousarray(np.stack([np.asarray(d[0], np.float32).reshape(16) for d in draws]))
    counts = np.array([np.asarray(d[1]).reshape(-1, 9).shape[0] for d in draws], np.int32)
    pos = _c(np.concatenate([np.asarray(d[1], np.float32).reshape(-1) for d in draws]), np.float32)
    out = np.zeros((H, W, 4), np.float32)
    lib.orc_voxel_view_f32(W, H, _f(_c(view, np.float32)), _f(_c(proj, np.float32)), len(draws), _f(Ms),
                           counts.ctypes.data, _f(pos), res, levels, arr, cube_size, view_dir, view_lod, _f(out))
    return out


def decode_rgba8(a: np.ndarray) -> np.ndarray:
    a = _c(a, np.uint32)
    out = np.zeros(a.shape + (4,), np.float32)
    load().orc_decode_rgba8(_u(a), _f(out), a.size)
    return out


def decode_f16(a: np.ndarray) -> np.ndarray:
    a = _c(a, np.uint16)
    out = np.zeros(a.shape, np.float32)
    load().orc_decode_f16(a.ctypes.data_as(C.POINTER(C.c_uint16)), _f(out), a.size)
    return out


def decode_bf16(a: np.ndarray) -> np.ndarray:
    a = _c(a, np.uint16)
    out = np.zeros(a.shape, np.float32)
    load().orc_decode_bf16(a.ctypes.data_as(C.POINTER(C.c_uint16)), _f(out), a.size)
    return out


def encode_f16(a: np.ndarray) -> np.ndarray:
    a = _c(a, np.float32)
    out = np.zeros(a.shape, np.uint16)
    load().orc_encode_f16(_f(a), out.ctypes.data_as(C.POINTER(C.c_uint16)), a.size)
    return out


def encode_bf16(a: np.ndarray) -> np.ndarray:
    a = _c(a, np.float32)
    out = np.zeros(a.shape, np.uint16)
    load().orc_encode_bf16(_f(a), out.ctypes.data_as(C.POINTER(C.c_uint16)), a.size)
    return out


def encode_rgba8(a: np.ndarray) -> np.ndarray:
    a = _c(a, np.float32)
    out = np.zeros(a.shape[:-1], np.uint32)
    load().orc_encode_rgba8(_f(a), _u(out), out.size)
    return out


def store_round(a: np.ndarray, dtype: int) -> np.ndarray:
    a = np.array(a, np.float32, copy=True, order="C")
    load().orc_store_round(_f(a), a.size, dtype)
    return a


def psnr(a: np.ndarray, b: np.ndarray, peak: float = 1.0) -> float:
    mse = float(np.mean((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2))
    return float("inf") if mse == 0 else 10.0 * np.log10(peak * peak / mse)


# ---- probes of single functions (pinning tests) -----------------------------------
def _probe_lib():
    lib = load()
    if not getattr(lib, "_probes", False):
        lib.orc_sample_lod_f64.argtypes = [C.c_int, C.c_int, C.POINTER(_DP), C.c_int, _DP, C.c_double, _DP]
        lib.orc_diffuse_dirs_f64.argtypes = [_DP, C.c_int, _DP]
        lib.orc_diffuse_dirs_f64.restype = C.c_int
        lib.orc_trace_cone_f32.argtypes = [C.c_int, C.c_int, C.POINTER(_FP), _FP, _FP, C.c_float, C.c_float, _FP]
        lib.orc_trace_cone_f32.restype = C.c_int64
        lib.orc_raster_coverage.argtypes = [_FP, _FP, C.c_int, C.c_int, C.POINTER(C.c_uint8)]
        lib.orc_dominant_axis.argtypes = [_FP]
        lib.orc_dominant_axis.restype = C.c_int
        lib._probes = True
    return lib


def _volume_ptrs(levels_list, dt):
    """levels_list[l][d] arrays (level 0: one array reused for all d)."""
    P = C.POINTER(C.c_double) if dt == np.float64 else C.POINTER(C.c_float)
    keep, ptrs = [], []
    for l, lv in enumerate(levels_list):
        for d in range(6):
            a = _c(lv if l == 0 else lv[d], dt)
            keep.append(a)
            ptrs.append(a.ctypes.data_as(P))
    return keep, (P * len(ptrs))(*ptrs)


def sample_lod(levels_list, map_index, p, lod):
    lib = _probe_lib()
    res = levels_list[0].shape[0]
    keep, arr = _volume_ptrs(levels_list, np.float64)
    pp = _c(p, np.float64)
    out = np.zeros(4, np.float64)
    lib.orc_sample_lod_f64(res, len(levels_list), arr, map_index, _d(pp), float(lod), _d(out))
    return out


def diffuse_dirs(n, cone_set=0):
    lib = _probe_lib()
    nn = _c(n, np.float64)
    out = np.zeros((9, 3), np.float64)
    k = lib.orc_diffuse_dirs_f64(_d(nn), cone_set, _d(out))
    return out[:k]


def trace_cone(levels_list, origin, direction, aperture, max_dist):
    lib = _probe_lib()
    res = levels_list[0].shape[0]
    keep, arr = _volume_ptrs(levels_list, np.float32)
    o = _c(origin, np.float32)
    d = _c(direction, np.float32)
    out = np.zeros(4, np.float32)
    steps = lib.orc_trace_cone_f32(res, len(levels_list), arr, _f(o), _f(d), aperture, max_dist, _f(out))
    return out, int(steps)


def raster_coverage(xw, yw, W, H):
    lib = _probe_lib()
    x = _c(xw, np.float32)
    y = _c(yw, np.float32)
    m = np.zeros((H, W), np.uint8)
    lib.orc_raster_coverage(_f(x), _f(y), W, H, m.ctypes.data_as(C.POINTER(C.c_uint8)))
    return m


def mlp_forward(params, x, bf16=False):
    lib = load()
    lib.orc_mlp_forward.argtypes = [_FP, _FP, C.c_int, C.c_int, _DP]
    p = _c(params, np.float32)
    xx = _c(x, np.float32)
    y = np.zeros((xx.shape[0], 3), np.float64)
    lib.orc_mlp_forward(_f(p), _f(xx), xx.shape[0], int(bf16), _d(y))
    return y


def mlp_loss_grad(params, x, albedo, target, bf16=False):
    lib = load()
    lib.orc_mlp_loss_grad.argtypes = [_FP, _FP, _FP, _FP, C.c_int, C.c_int, _DP]
    lib.orc_mlp_loss_grad.restype = C.c_double
    p = _c(params, np.float32)
    xx = _c(x, np.float32)
    a = _c(albedo, np.float32)
    t = _c(target, np.float32)
    g = np.zeros(p.size, np.float64)
    loss = lib.orc_mlp_loss_grad(_f(p), _f(xx), _f(a), _f(t), xx.shape[0], int(bf16), _d(g))
    return float(loss), g


def dominant_axis(tri):
    lib = _probe_lib()
    t = _c(tri, np.float32).reshape(9)
    return int(lib.orc_dominant_axis(_f(t)))
