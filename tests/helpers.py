"""Shared helpers of the parity tests: scenes, oracle inputs, comparisons."""
from __future__ import annotations

import numpy as np

from voxel_cone_tracing_amd import _lib as L
from voxel_cone_tracing_amd.scene import identity, look_at, perspective, soup_materials_fast


def default_material() -> L.Material:
    """The material a context starts with (renderer.cpp:85-98)."""
    m = L.Material()
    m.ambient[0] = 1
    m.ambient[2] = 1
    m.diffuse[0] = 1
    m.diffuse[2] = 1
    m.shininess = 1
    m.ior = 1
    m.dissolve = 1
    m.illum = 0
    return m


def material_table(extra) -> np.ndarray:
    """Full material table as (n, 32) float32 records: default + extra."""
    import oracle

    head = oracle.material_bytes([default_material()])
    if isinstance(extra, np.ndarray):
        return np.concatenate([head, extra.reshape(-1, 32)])
    return np.concatenate([head, oracle.material_bytes(extra)])


class SceneSpec:
    """Geometry + materials + lights + camera of one test scene."""

    def __init__(self, kind: str, seed: int = 1, ntri: int = 2000, res: int = 64, W: int = 96, H: int = 64,
                 cube_size: float | None = None, extent: float = 1.0, mu: float = -4.0, glass: bool = False,
                 phases=(False, True, True, False), eye=None, target=None, fovy_deg: float = 50.0,
                 model: str | None = None):
        self.kind, self.res, self.W, self.H = kind, res, W, H
        self.phases = tuple(phases)
        if kind == "cornell":
            pos, nrm, tm, mats = L.cornell(seed)
            if glass:
                # the second box transmissive and glossy (illum 4, dissolve <= 0.1: the
                # refraction override and refracted highlight of the trace), the first
                # one emissive
                g = mats[4]
                g.illum, g.dissolve, g.ior, g.shininess = 4, 0.05, 1.5, 64.0
                for k, (t, sp) in enumerate(zip((0.9, 0.95, 0.9), (0.5, 0.5, 0.5))):
                    g.transmittance[k], g.specular[k] = t, sp
                for k, e in enumerate((0.2, 0.1, 0.05)):
                    mats[3].emission[k] = e
            self.pos, self.nrm = pos, nrm
            self.tri_mat = (tm + 1).astype(np.uint32)  # after the default material
            self.mat_records = material_table(mats)
            self.extra = mats
            self.cube_size = 0.55 if cube_size is None else cube_size
            self.lights = [((0.0, 0.45, 0.0), (1.0, 1.0, 1.0), 1.0, L.LIGHT_POINT)]
            deye, dtgt = (0.0, 0.0, 1.55), (0.0, 0.0, 0.0)
            self.cone_set = L.CONES_5_WIDE
        else:
            pos, nrm, ar = L.soup(seed, ntri, extent, mu)
            self.pos, self.nrm = pos, nrm
            rec = soup_materials_fast(ar)
            self.tri_mat = (np.arange(ntri, dtype=np.uint32) + 1).astype(np.uint32)
            self.mat_records = material_table(rec)
            self.extra = rec
            self.cube_size = 1.1 if cube_size is None else cube_size
            self.lights = [((0.4, 1.0, 0.6), (1.0, 1.0, 1.0), 1.0, L.LIGHT_DIRECTIONAL)]
            deye, dtgt = (0.0, 0.3, 3.2), (0.0, 0.0, 0.0)
            self.cone_set = L.CONES_REF9
        self.M = identity() if model is None else model_matrix(model)
        if eye is None:
            eye, target = deye, dtgt
        self.view = look_at(eye, target, (0, 1, 0))
        self.proj = perspective(np.radians(fovy_deg), W / H, 0.1, 100.0)

    @property
    def cam(self):
        # camera_position = column 3 of the view matrix (renderer.cpp:279)
        return self.view.reshape(16)[12:15].copy()

    def build(self, ctx, dtype=L.RGBA8, average=None, levels=7):
        levels = min(levels, self.res.bit_length())  # at most log2(res) + 1 levels
        ctx.set_grid(self.res, levels, dtype, average)
        ctx.set_grid_size(self.cube_size)
        ctx.set_cone_set(self.cone_set)
        ctx.set_phases(*self.phases)
        ctx.set_voxel_view(7, 0.0)
        if isinstance(self.extra, np.ndarray):
            from voxel_cone_tracing_amd.scene import C_add_material_records

            first = C_add_material_records(ctx, self.extra)
        else:
            first = ctx.add_materials(self.extra)
        assert first == 1
        self.model = ctx.model_create(self.pos, self.nrm, self.tri_mat)
        ctx.set_model_transform(self.model, self.M)
        ctx.set_camera(self.view, self.proj)
        ctx.clear_queues()
        ctx.queue_model(self.model)
        for (p, c, i, k) in self.lights:
            ctx.queue_light(p, c, i, k)
        return self.model


def model_matrix(name: str) -> np.ndarray:
    """Column-major model matrices of the transform tests: "rot_scale" is a
    translation, two rotations and a non-uniform scale (so the normal matrix,
    mat3(transpose(inverse(model))), differs from the model's own 3x3)."""
    if name != "rot_scale":
        raise ValueError(name)

    def rot(axis, ang):
        c, s = np.cos(ang), np.sin(ang)
        R = np.eye(4)
        i, j = [(1, 2), (2, 0), (0, 1)][axis]
        R[i, i], R[j, j], R[i, j], R[j, i] = c, c, -s, s
        return R

    T = np.eye(4)
    T[:3, 3] = (0.05, -0.04, 0.03)
    S = np.diag([0.85, 0.6, 0.95, 1.0])
    return np.ascontiguousarray((T @ rot(1, 0.5) @ rot(0, -0.35) @ S).T.astype(np.float32))


def decode_level(arr: np.ndarray, dtype: int) -> np.ndarray:
    """Stored level (as downloaded) -> float32 (..., 4)."""
    import oracle

    if dtype == L.RGBA8:
        return oracle.decode_rgba8(arr)
    if dtype == L.FP16:
        return oracle.decode_f16(arr)
    if dtype == L.BF16:
        return oracle.decode_bf16(arr)
    return np.asarray(arr, np.float32)


def encode_level(f: np.ndarray, dtype: int) -> np.ndarray:
    import oracle

    if dtype == L.RGBA8:
        return oracle.encode_rgba8(f)
    if dtype == L.FP16:
        return oracle.encode_f16(f)
    if dtype == L.BF16:
        return oracle.encode_bf16(f)
    return np.asarray(f, np.float32)


def bits(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a)
    if a.dtype == np.float32:
        return a.view(np.uint32)
    return a
