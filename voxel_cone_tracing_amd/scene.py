"""Camera math and the procedural workloads.

look_at / perspective follow glm's conventions (column-major, right-handed,
clip z in [-w, w]) because the reference builds its matrices with
glm::lookAt / glm::perspective (camera.h:23-60, main.cpp:107-109).

Workloads (sizes of the benchmark configurations):
  cornell64   procedural Cornell box, 64^3 grid, 512x512, 5 wide cones
  soup128     200k-triangle soup, 128^3, 1920x1080, 9 cones + specular
  soup256     1M-triangle soup, 256^3, 2560x1440, 9 cones + specular
  soup512     5M-triangle soup, 512^3, 3840x2160, 16 cones + specular
Soups: centres U[-1,1]^3, edge LogNormal(mu=-4, sigma=0.5), uniform random
orientation, per-triangle albedo U[0,1]^3, roughness U[0.05,0.5] mapped to
shininess = 2/r^2 - 2, specular colour = albedo / 2, one directional light.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L


def look_at(eye, center, up) -> np.ndarray:
    eye, center, up = (np.asarray(v, np.float64) for v in (eye, center, up))
    f = center - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    m = np.eye(4)
    m[0, :3] = s
    m[1, :3] = u
    m[2, :3] = -f
    m[0, 3] = -s.dot(eye)
    m[1, 3] = -u.dot(eye)
    m[2, 3] = f.dot(eye)
    # column-major storage = transpose of the row-major matrix
    return np.ascontiguousarray(m.T.astype(np.float32))


def perspective(fovy_rad: float, aspect: float, near: float, far: float) -> np.ndarray:
    t = math.tan(fovy_rad / 2.0)
    m = np.zeros((4, 4))
    m[0, 0] = 1.0 / (aspect * t)
    m[1, 1] = 1.0 / t
    m[2, 2] = -(far + near) / (far - near)
    m[3, 2] = -1.0
    m[2, 3] = -(2.0 * far * near) / (far - near)
    return np.ascontiguousarray(m.T.astype(np.float32))


def identity() -> np.ndarray:
    return np.eye(4, dtype=np.float32)


@dataclass
class Workload:
    name: str
    res: int
    width: int
    height: int
    cone_set: int
    cube_size: float
    ntri: int
    levels: int = 7
    seed: int = 1
    eye: tuple = (0.0, 0.3, 3.2)
    target: tuple = (0.0, 0.0, 0.0)
    fovy: float = math.radians(50.0)
    lights: list = field(default_factory=list)


WORKLOADS = {
    "cornell64": Workload("cornell64", 64, 512, 512, L.CONES_5_WIDE, 0.55, 34,
                          eye=(0.0, 0.0, 1.55), target=(0.0, 0.0, 0.0), fovy=math.radians(45.0),
                          lights=[((0.0, 0.45, 0.0), (1.0, 1.0, 1.0), 1.0, L.LIGHT_POINT)]),
    "soup128": Workload("soup128", 128, 1920, 1080, L.CONES_REF9, 1.1, 200_000,
                        lights=[((0.4, 1.0, 0.6), (1.0, 1.0, 1.0), 1.0, L.LIGHT_DIRECTIONAL)]),
    "soup256": Workload("soup256", 256, 2560, 1440, L.CONES_REF9, 1.1, 1_000_000,
                        lights=[((0.4, 1.0, 0.6), (1.0, 1.0, 1.0), 1.0, L.LIGHT_DIRECTIONAL)]),
    "soup512": Workload("soup512", 512, 3840, 2160, L.CONES_16, 1.1, 5_000_000,
                        lights=[((0.4, 1.0, 0.6), (1.0, 1.0, 1.0), 1.0, L.LIGHT_DIRECTIONAL)]),
}


def soup_materials(albedo_rough: np.ndarray):
    """One material per triangle: Kd = albedo, Ks = albedo / 2, shininess from roughness."""
    mats = []
    for a in albedo_rough:
        r = float(a[3])
        shin = 2.0 / (r * r) - 2.0
        mats.append(L.make_material(diffuse=a[:3], specular=a[:3] * 0.5, shininess=shin))
    return mats


def soup_materials_fast(albedo_rough: np.ndarray) -> "np.ndarray":
    """Same as soup_materials, built as a structured byte array (fast for 1M+)."""
    n = albedo_rough.shape[0]
    rec = np.zeros((n, 32), np.float32)  # 128-byte material_data_t records
    a = albedo_rough[:, :3]
    rec[:, 0:3] = a           # ambient
    rec[:, 4:7] = a           # diffuse
    rec[:, 8:11] = a * 0.5    # specular
    rec[:, 12:15] = 1.0       # transmittance
    r = albedo_rough[:, 3].astype(np.float64)
    rec[:, 19] = (2.0 / (r * r) - 2.0).astype(np.float32)  # shininess
    rec[:, 20] = 1.0          # ior
    rec[:, 21] = 1.0          # dissolve
    rec[:, 22] = np.array([2], np.int32).view(np.float32)[0]  # illum = 2
    return rec


class Scene:
    """A workload built into a native context (geometry resident on the GPU)."""

    def __init__(self, ctx, wl: Workload, dtype: int = L.FP16, average: int | None = None):
        self.ctx, self.wl = ctx, wl
        ctx.set_grid(wl.res, wl.levels, dtype, average)
        ctx.set_grid_size(wl.cube_size)
        ctx.set_cone_set(wl.cone_set)
        ctx.set_phases(False, True, True, False)  # indirect only
        ctx.set_voxel_view(7, 0.0)
        if wl.name.startswith("cornell"):
            pos, nrm, tm, mats = L.cornell(wl.seed)
            base = ctx.add_materials(mats)
            self.model = ctx.model_create(pos, nrm, tm + base)
            self.pos, self.nrm, self.tri_mat, self.mats = pos, nrm, tm + base, mats
        else:
            pos, nrm, ar = L.soup(wl.seed, wl.ntri)
            rec = soup_materials_fast(ar)
            first = C_add_material_records(ctx, rec)
            tm = (np.arange(wl.ntri, dtype=np.uint32) + first).astype(np.uint32)
            self.model = ctx.model_create(pos, nrm, tm)
            self.pos, self.nrm, self.tri_mat, self.albedo_rough = pos, nrm, tm, ar
        self.view = look_at(wl.eye, wl.target, (0, 1, 0))
        self.proj = perspective(wl.fovy, wl.width / wl.height, 0.1, 100.0)
        ctx.set_camera(self.view, self.proj)

    def queue(self) -> None:
        ctx = self.ctx
        ctx.clear_queues()
        ctx.queue_model(self.model)
        for (p, c, i, k) in self.wl.lights:
            ctx.queue_light(p, c, i, k)


def C_add_material_records(ctx, rec: np.ndarray) -> int:
    import ctypes as C

    rec = np.ascontiguousarray(rec, np.float32)
    first = C.c_uint32()
    L.check(ctx.lib.vct_add_materials(ctx.h, rec.ctypes.data_as(C.POINTER(L.Material)), rec.shape[0],
                                      C.byref(first)))
    return first.value
