"""ctypes binding of libvct.so (the C ABI declared in include/vct.h).

This is plumbing for Python callers (tests, bench.py): every entry point is a
thin call into the native library.  There is no Python fallback: if libvct.so
is missing or cannot be loaded, load() raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libvct.so"

VCT_OK = 0
VCT_ERR_INVALID_ARG = -1
VCT_ERR_BAD_HANDLE = -2
VCT_ERR_OUT_OF_MEMORY = -3
VCT_ERR_HIP = -4
VCT_ERR_NO_DEVICE = -5
VCT_ERR_STATE = -6
VCT_ERR_LIMIT = -7
VCT_ERR_NOT_FOUND = -8
VCT_INVALID_ID = 0xFFFFFFFF

RGBA8, FP16, BF16, FP32 = 0, 1, 2, 3
AVG_REFERENCE, AVG_EXACT = 0, 1
CONES_REF9, CONES_5_WIDE, CONES_16 = 0, 1, 2
LIGHT_POINT, LIGHT_DIRECTIONAL = 0, 1
MLP_IN, MLP_HIDDEN, MLP_OUT = 34, 64, 3
MLP_NOALLOC_MAX_ROWS = 262144
MLP_PARAMS = MLP_IN * MLP_HIDDEN + MLP_HIDDEN + MLP_HIDDEN * MLP_HIDDEN + MLP_HIDDEN + MLP_HIDDEN * MLP_OUT + MLP_OUT


class Material(C.Structure):
    """material_data_t (renderer.h:94-120), 128 bytes."""

    _pack_ = 1
    _fields_ = [
        ("ambient", C.c_float * 4),
        ("diffuse", C.c_float * 4),
        ("specular", C.c_float * 4),
        ("transmittance", C.c_float * 4),
        ("emission", C.c_float * 3),
        ("shininess", C.c_float),
        ("ior", C.c_float),
        ("dissolve", C.c_float),
        ("illum", C.c_int32),
        ("roughness", C.c_float),
        ("metallic", C.c_float),
        ("sheen", C.c_float),
        ("clearcoat_thickness", C.c_float),
        ("clearcoat_roughness", C.c_float),
        ("anisotropy", C.c_float),
        ("anisotropy_rotation", C.c_float),
        ("pad", C.c_float * 2),
    ]


assert C.sizeof(Material) == 128


class Light(C.Structure):
    _fields_ = [
        ("position", C.c_float * 3),
        ("color", C.c_float * 3),
        ("intensity", C.c_float),
        ("kind", C.c_int32),
    ]


class VctError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"vct error {code}: {msg}")
        self.code = code


_lib = None

_P = C.c_void_p
_I = C.c_int
_U = C.c_uint32
_F = C.c_float
_SZ = C.c_size_t
_FP = C.POINTER(C.c_float)
_UP = C.POINTER(C.c_uint32)
_U64P = C.POINTER(C.c_uint64)

# name -> (restype, argtypes)
PROTOTYPES = {
    "vct_last_error": (C.c_char_p, []),
    "vct_device_count": (_I, [C.POINTER(C.c_int)]),
    "vct_create": (_I, [_I, _I, C.POINTER(_P)]),
    "vct_destroy": (_I, [_P]),
    "vct_sync": (_I, [_P]),
    "vct_set_viewport": (_I, [_P, _I, _I]),
    "vct_set_grid_resolution": (_I, [_P, C.c_uint]),
    "vct_set_grid_format": (_I, [_P, C.c_uint, _I, _I, _I]),
    "vct_set_grid_size": (_I, [_P, _F]),
    "vct_get_grid": (_I, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "vct_set_camera_transform": (_I, [_P, _FP, _FP]),
    "vct_add_material": (_I, [_P, C.POINTER(Material), _UP]),
    "vct_add_materials": (_I, [_P, C.POINTER(Material), _U, _UP]),
    "vct_upload_material_data": (_I, [_P, C.POINTER(Material), _U]),
    "vct_get_material": (_I, [_P, _U, C.POINTER(Material)]),
    "vct_model_create": (_I, [_P, _FP, _FP, _UP, _U, _UP]),
    "vct_load_model": (_I, [_P, C.c_char_p, _UP]),
    "vct_set_model_material": (_I, [_P, _U, _U]),
    "vct_set_model_transform": (_I, [_P, _U, _FP]),
    "vct_get_model_dimensions": (_I, [_P, _U, _FP]),
    "vct_model_triangle_count": (_I, [_P, _U, _UP]),
    "vct_queue_model": (_I, [_P, _U]),
    "vct_queue_point_light": (_I, [_P, C.POINTER(Light)]),
    "vct_clear_queues": (_I, [_P]),
    "vct_set_rendering_phases": (_I, [_P, _I, _I, _I, _I]),
    "vct_set_voxel_view_dir": (_I, [_P, _I, _F]),
    "vct_set_cone_set": (_I, [_P, _I]),
    "vct_voxelize": (_I, [_P]),
    "vct_mip_aniso": (_I, [_P]),
    "vct_gbuffer": (_I, [_P]),
    "vct_trace": (_I, [_P]),
    "vct_render": (_I, [_P]),
    "vct_volume_level_bytes": (_SZ, [_P, _I]),
    "vct_volume_download": (_I, [_P, _I, _I, _P, _SZ]),
    "vct_volume_upload": (_I, [_P, _I, _I, _P, _SZ]),
    "vct_gbuffer_download": (_I, [_P, _FP, _FP, _SZ]),
    "vct_gbuffer_upload": (_I, [_P, _FP, _FP, _SZ]),
    "vct_image_download": (_I, [_P, _FP, _SZ]),
    "vct_get_pass_times": (_I, [_P, _FP, _I]),
    "vct_set_pass_timing": (_I, [_P, _U]),
    "vct_get_voxelize_stats": (_I, [_P, _U64P, _U64P]),
    "vct_trace_count_steps": (_I, [_P, _U64P]),
    "vct_tex3d_create": (_I, [_P, _I, _I, _I, _I, _I, _UP]),
    "vct_tex3d_destroy": (_I, [_P, _U]),
    "vct_tex3d_activate": (_I, [_P, _U, _U]),
    "vct_tex3d_clear": (_I, [_P, _U, _FP]),
    "vct_tex3d_mip": (_I, [_P, _U]),
    "vct_tex3d_upload": (_I, [_P, _U, _I, _P, _SZ]),
    "vct_tex3d_download": (_I, [_P, _U, _I, _P, _SZ]),
    "vct_mlp_init": (_I, [_P, C.c_uint64]),
    "vct_mlp_set_params": (_I, [_P, _FP, _U]),
    "vct_mlp_get_params": (_I, [_P, _FP, _U]),
    "vct_mlp_forward": (_I, [_P, _FP, _U, _FP]),
    "vct_mlp_forward_noalloc": (_I, [_P, _P, _U, _P, _P]),
    "vct_trace_load_peak": (_I, [_P, C.POINTER(C.c_double)]),
    "vct_mlp_loss_grad": (_I, [_P, _FP, _FP, _FP, _U, _FP, _FP]),
    "vct_upsample_collect": (_I, [_P, _U, C.c_uint64]),
    "vct_upsample_clear_samples": (_I, [_P]),
    "vct_upsample_set_samples": (_I, [_P, _FP, _FP, _FP, _U]),
    "vct_upsample_sample_count": (_I, [_P, _U64P]),
    "vct_upsample_train": (_I, [_P, _U, _U, _F, C.c_uint64, _FP]),
    "vct_render_learned": (_I, [_P]),
    "vct_upsample_lowres_download": (_I, [_P, _FP, _SZ]),
    "vct_upsample_set_fallback": (_I, [_P, _F]),
    "vct_upsample_fallback_stats": (_I, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "vct_upsample_traced_pixels": (_I, [_P, C.POINTER(C.c_uint32)]),
    "vct_upsample_set_refine": (_I, [_P, _I]),
    "vct_upsample_stats": (_I, [_P, C.POINTER(C.c_uint32)]),
    "vct_obj_open": (_I, [C.c_char_p, C.POINTER(C.c_void_p), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "vct_obj_read": (_I, [C.c_void_p, _FP, _FP, C.POINTER(C.c_int32), C.c_void_p, _FP]),
    "vct_obj_close": (None, [C.c_void_p]),
    "vct_scene_soup": (_I, [C.c_uint64, _U, _F, _F, _F, _F, _F, _FP, _FP, _FP]),
    "vct_scene_cornell": (_I, [C.c_uint64, _FP, _FP, _UP, C.POINTER(Material)]),
}


def load() -> C.CDLL:
    """Load libvct.so (raises if it is missing: there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: run `python -m voxel_cone_tracing_amd.build` first")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != VCT_OK:
        msg = _lib.vct_last_error().decode() if _lib is not None else ""
        raise VctError(rc, msg)


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_FP)


def uptr(a: np.ndarray):
    assert a.dtype == np.uint32 and a.flags.c_contiguous
    return a.ctypes.data_as(_UP)


def vptr(a: np.ndarray):
    assert a.flags.c_contiguous
    return C.c_void_p(a.ctypes.data)


def device_count() -> int:
    lib = load()
    n = C.c_int(0)
    check(lib.vct_device_count(C.byref(n)))
    return n.value


def make_material(diffuse=(1, 1, 1), specular=(0, 0, 0), emission=(0, 0, 0), transmittance=(1, 1, 1),
                  shininess=1.0, ior=1.0, dissolve=1.0, illum=2) -> Material:
    m = Material()
    for k in range(3):
        m.ambient[k] = diffuse[k]
        m.diffuse[k] = diffuse[k]
        m.specular[k] = specular[k]
        m.emission[k] = emission[k]
        m.transmittance[k] = transmittance[k]
    m.shininess = shininess
    m.ior = ior
    m.dissolve = dissolve
    m.illum = illum
    return m


def materials_array(mats) -> "C.Array":
    arr = (Material * len(mats))()
    for i, m in enumerate(mats):
        arr[i] = m
    return arr


def soup(seed: int, ntri: int, extent: float = 1.0, mu: float = -4.0, sigma: float = 0.5,
         rmin: float = 0.05, rmax: float = 0.5):
    """Random triangle soup (host generator in libvct.so): pos, nrm (ntri x 9), albedo_rough (ntri x 4)."""
    lib = load()
    pos = np.zeros((ntri, 9), np.float32)
    nrm = np.zeros((ntri, 9), np.float32)
    ar = np.zeros((ntri, 4), np.float32)
    check(lib.vct_scene_soup(seed, ntri, extent, mu, sigma, rmin, rmax, fptr(pos), fptr(nrm), fptr(ar)))
    return pos, nrm, ar


def read_obj(path: str):
    """Parses an OBJ/MTL file as load_model does (no GPU needed): positions and
    normals (n x 9), per-triangle local material index (-1: none), the list of
    Materials and the reference's model dimensions."""
    lib = load()
    h = C.c_void_p()
    nt, nm = C.c_uint32(), C.c_uint32()
    check(lib.vct_obj_open(str(path).encode(), C.byref(h), C.byref(nt), C.byref(nm)))
    try:
        pos = np.zeros((nt.value, 9), np.float32)
        nrm = np.zeros((nt.value, 9), np.float32)
        tm = np.zeros(nt.value, np.int32)
        mats = (Material * max(1, nm.value))()
        dims = np.zeros(3, np.float32)
        check(lib.vct_obj_read(h, fptr(pos), fptr(nrm), tm.ctypes.data_as(C.POINTER(C.c_int32)), mats, fptr(dims)))
    finally:
        lib.vct_obj_close(h)
    return pos, nrm, tm, [mats[i] for i in range(nm.value)], dims


def cornell(seed: int):
    """Procedural Cornell box: pos, nrm (34 x 9), tri_mat (34,), list of 5 Materials."""
    lib = load()
    pos = np.zeros((34, 9), np.float32)
    nrm = np.zeros((34, 9), np.float32)
    tm = np.zeros(34, np.uint32)
    mats = (Material * 5)()
    check(lib.vct_scene_cornell(seed, fptr(pos), fptr(nrm), uptr(tm), mats))
    return pos, nrm, tm, [mats[i] for i in range(5)]
