"""Python handle over a native vct context (include/vct.h).

Mirrors the reference Renderer's call sequence (renderer.h:122-155) for
Python callers.  All work happens in libvct.so on the GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


class Context:
    def __init__(self, width: int, height: int):
        self.lib = L.load()
        h = C.c_void_p()
        L.check(self.lib.vct_create(width, height, C.byref(h)))
        self.h = h
        self.width, self.height = width, height

    # -- lifetime ---------------------------------------------------------------
    def close(self) -> None:
        if self.h:
            self.lib.vct_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self) -> None:
        L.check(self.lib.vct_sync(self.h))

    # -- configuration ------------------------------------------------------------
    def set_viewport(self, w: int, h: int) -> None:
        L.check(self.lib.vct_set_viewport(self.h, w, h))
        self.width, self.height = w, h

    def set_grid(self, res: int, levels: int = 7, dtype: int = L.RGBA8, average: int | None = None) -> None:
        if average is None:
            average = L.AVG_REFERENCE if dtype == L.RGBA8 else L.AVG_EXACT
        L.check(self.lib.vct_set_grid_format(self.h, res, levels, dtype, average))

    def grid(self):
        r, l, d, a = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        L.check(self.lib.vct_get_grid(self.h, C.byref(r), C.byref(l), C.byref(d), C.byref(a)))
        return r.value, l.value, d.value, a.value

    def set_grid_size(self, s: float) -> None:
        L.check(self.lib.vct_set_grid_size(self.h, s))

    def set_camera(self, view: np.ndarray, proj: np.ndarray) -> None:
        v = np.ascontiguousarray(view, np.float32).reshape(16)
        p = np.ascontiguousarray(proj, np.float32).reshape(16)
        L.check(self.lib.vct_set_camera_transform(self.h, L.fptr(v), L.fptr(p)))

    def set_phases(self, direct: bool, diffuse: bool, specular: bool, shadow: bool) -> None:
        L.check(self.lib.vct_set_rendering_phases(self.h, int(direct), int(diffuse), int(specular), int(shadow)))

    def set_voxel_view(self, d: int, lod: float) -> None:
        L.check(self.lib.vct_set_voxel_view_dir(self.h, d, lod))

    def set_cone_set(self, s: int) -> None:
        L.check(self.lib.vct_set_cone_set(self.h, s))

    # -- scene --------------------------------------------------------------------
    def add_materials(self, mats) -> int:
        arr = L.materials_array(mats)
        first = C.c_uint32()
        L.check(self.lib.vct_add_materials(self.h, arr, len(mats), C.byref(first)))
        return first.value

    def model_create(self, pos: np.ndarray, nrm: np.ndarray, tri_mat: np.ndarray | None = None) -> int:
        pos = np.ascontiguousarray(pos, np.float32)
        nrm = np.ascontiguousarray(nrm, np.float32)
        n = pos.reshape(-1, 9).shape[0]
        out = C.c_uint32()
        tm = None
        if tri_mat is not None:
            tri_mat = np.ascontiguousarray(tri_mat, np.uint32)
            tm = L.uptr(tri_mat)
        L.check(self.lib.vct_model_create(self.h, L.fptr(pos), L.fptr(nrm), tm, n, C.byref(out)))
        return out.value

    def load_model(self, path: str) -> int:
        out = C.c_uint32()
        L.check(self.lib.vct_load_model(self.h, path.encode(), C.byref(out)))
        return out.value

    def set_model_transform(self, model: int, m: np.ndarray) -> None:
        a = np.ascontiguousarray(m, np.float32).reshape(16)
        L.check(self.lib.vct_set_model_transform(self.h, model, L.fptr(a)))

    def set_model_material(self, material: int, model: int) -> None:
        """Every triangle of the model takes the material (renderer.cpp:225-233)."""
        L.check(self.lib.vct_set_model_material(self.h, material, model))

    def model_dimensions(self, model: int) -> np.ndarray:
        d = np.zeros(3, np.float32)
        L.check(self.lib.vct_get_model_dimensions(self.h, model, L.fptr(d)))
        return d

    def queue_model(self, model: int) -> None:
        L.check(self.lib.vct_queue_model(self.h, model))

    def queue_light(self, position, color=(1, 1, 1), intensity=1.0, kind=L.LIGHT_POINT) -> None:
        li = L.Light()
        for k in range(3):
            li.position[k] = position[k]
            li.color[k] = color[k]
        li.intensity = intensity
        li.kind = kind
        L.check(self.lib.vct_queue_point_light(self.h, C.byref(li)))

    def clear_queues(self) -> None:
        L.check(self.lib.vct_clear_queues(self.h))

    # -- passes -------------------------------------------------------------------
    def voxelize(self) -> None:
        L.check(self.lib.vct_voxelize(self.h))

    def mip(self) -> None:
        L.check(self.lib.vct_mip_aniso(self.h))

    def gbuffer(self) -> None:
        L.check(self.lib.vct_gbuffer(self.h))

    def trace(self) -> None:
        L.check(self.lib.vct_trace(self.h))

    def render(self) -> None:
        L.check(self.lib.vct_render(self.h))

    # -- read back ----------------------------------------------------------------
    def download_level(self, level: int, d: int = 0) -> np.ndarray:
        res, _, dt, _ = self.grid()
        s = res >> level
        nbytes = self.lib.vct_volume_level_bytes(self.h, level)
        if dt == L.RGBA8:
            a = np.zeros((s, s, s), np.uint32)
        elif dt == L.FP32:
            a = np.zeros((s, s, s, 4), np.float32)
        else:
            a = np.zeros((s, s, s, 4), np.uint16)
        assert a.nbytes == nbytes
        L.check(self.lib.vct_volume_download(self.h, level, d, L.vptr(a), nbytes))
        return a

    def upload_level(self, level: int, d: int, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        L.check(self.lib.vct_volume_upload(self.h, level, d, L.vptr(a), a.nbytes))

    def download_gbuffer(self):
        n = self.width * self.height
        p = np.zeros((self.height, self.width, 4), np.float32)
        q = np.zeros((self.height, self.width, 4), np.float32)
        L.check(self.lib.vct_gbuffer_download(self.h, L.fptr(p), L.fptr(q), n))
        return p, q

    def upload_gbuffer(self, p: np.ndarray, q: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        q = np.ascontiguousarray(q, np.float32)
        L.check(self.lib.vct_gbuffer_upload(self.h, L.fptr(p), L.fptr(q), self.width * self.height))

    def download_image(self) -> np.ndarray:
        img = np.zeros((self.height, self.width, 4), np.float32)
        L.check(self.lib.vct_image_download(self.h, L.fptr(img), self.width * self.height))
        return img

    def pass_times(self) -> np.ndarray:
        """[voxelize, mip, gbuffer, trace, upsample, fallback] of the last frame (ms)."""
        t = np.zeros(6, np.float32)
        L.check(self.lib.vct_get_pass_times(self.h, L.fptr(t), t.shape[0]))
        return t

    def set_pass_timing(self, mask: int) -> None:
        """Passes that record timing events (bit k: pass k of pass_times; 0x3f all, 0 none)."""
        L.check(self.lib.vct_set_pass_timing(self.h, mask))

    # -- learned upsampler --------------------------------------------------------
    def mlp_init(self, seed: int = 1) -> None:
        L.check(self.lib.vct_mlp_init(self.h, seed))

    def mlp_params(self) -> np.ndarray:
        p = np.zeros(L.MLP_PARAMS, np.float32)
        L.check(self.lib.vct_mlp_get_params(self.h, L.fptr(p), p.size))
        return p

    def mlp_set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        L.check(self.lib.vct_mlp_set_params(self.h, L.fptr(p), p.size))

    def mlp_forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        n = x.shape[0]
        y = np.zeros((n, 3), np.float32)
        L.check(self.lib.vct_mlp_forward(self.h, L.fptr(x), n, L.fptr(y)))
        return y

    def mlp_loss_grad(self, x, albedo, target):
        x = np.ascontiguousarray(x, np.float32)
        a = np.ascontiguousarray(albedo, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        g = np.zeros(L.MLP_PARAMS, np.float32)
        loss = C.c_float()
        L.check(self.lib.vct_mlp_loss_grad(self.h, L.fptr(x), L.fptr(a), L.fptr(t), x.shape[0], L.fptr(g),
                                           C.byref(loss)))
        return float(loss.value), g

    def upsample_collect(self, n: int, seed: int) -> None:
        L.check(self.lib.vct_upsample_collect(self.h, n, seed))

    def upsample_set_samples(self, x, albedo, target) -> None:
        x = np.ascontiguousarray(x, np.float32)
        a = np.ascontiguousarray(albedo, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        L.check(self.lib.vct_upsample_set_samples(self.h, L.fptr(x), L.fptr(a), L.fptr(t), x.shape[0]))

    def upsample_clear(self) -> None:
        L.check(self.lib.vct_upsample_clear_samples(self.h))

    def upsample_count(self) -> int:
        n = C.c_uint64()
        L.check(self.lib.vct_upsample_sample_count(self.h, C.byref(n)))
        return n.value

    def upsample_train(self, steps: int, batch: int, lr: float, seed: int = 0) -> float:
        loss = C.c_float()
        L.check(self.lib.vct_upsample_train(self.h, steps, batch, lr, seed, C.byref(loss)))
        return float(loss.value)

    def render_learned(self) -> None:
        L.check(self.lib.vct_render_learned(self.h))

    def set_fallback(self, min_tile_psnr_db: float) -> None:
        L.check(self.lib.vct_upsample_set_fallback(self.h, float(min_tile_psnr_db)))

    def fallback_stats(self):
        """(tiles, tiles traced again) of the last learned frame."""
        t, f = C.c_uint32(), C.c_uint32()
        L.check(self.lib.vct_upsample_fallback_stats(self.h, C.byref(t), C.byref(f)))
        return t.value, f.value

    def set_refine(self, half_resolution: bool) -> None:
        L.check(self.lib.vct_upsample_set_refine(self.h, int(bool(half_resolution))))

    def upsample_stats(self) -> dict:
        """Counters of the last learned frame."""
        a = (C.c_uint32 * 6)()
        L.check(self.lib.vct_upsample_stats(self.h, a))
        keys = ["tiles", "tiles_refined", "tiles_full", "pixels_traced_quarter", "pixels_traced_half", "probes"]
        return dict(zip(keys, [int(x) for x in a]))

    def traced_pixels(self) -> int:
        """Pixels of the last learned frame traced at full resolution (no
        quarter-resolution sample of their surface nearby)."""
        n = C.c_uint32()
        L.check(self.lib.vct_upsample_traced_pixels(self.h, C.byref(n)))
        return n.value

    def download_lowres(self) -> np.ndarray:
        img = np.zeros((self.height // 4, self.width // 4, 4), np.float32)
        L.check(self.lib.vct_upsample_lowres_download(self.h, L.fptr(img), img.shape[0] * img.shape[1]))
        return img

    def trace_steps(self) -> int:
        n = C.c_uint64()
        L.check(self.lib.vct_trace_count_steps(self.h, C.byref(n)))
        return n.value

    def load_peak_gbps(self) -> float:
        """Ceiling of the marcher's texel loads (L1-hit 16-byte loads, GB/s)."""
        g = C.c_double()
        L.check(self.lib.vct_trace_load_peak(self.h, C.byref(g)))
        return g.value

    def voxelize_stats(self):
        f, c = C.c_uint64(), C.c_uint64()
        L.check(self.lib.vct_get_voxelize_stats(self.h, C.byref(f), C.byref(c)))
        return f.value, c.value
