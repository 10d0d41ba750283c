"""GPU kernels against the reference's own shaders (llvmpipe fixtures).

tests/golden/glref_*.npz hold what the reference GLSL (voxelize, mipmap.comp,
voxel_cone_tracing.frag) computed under Mesa llvmpipe for two seeded scenes
(tests/golden/make_glref_fixtures.py).  The GPU runs the same scenes through
the C ABI.  The running average is order dependent, so voxels reached by
several triangles may differ by the rounding of a different order (<= 2 LSB);
everything else is compared exactly or within the stated tolerance.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from voxel_cone_tracing_amd import _lib as L
from voxel_cone_tracing_amd.context import Context

from helpers import SceneSpec

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"
SCENES = {
    "cornell64": dict(kind="cornell", res=64, W=128, H=96),
    "soup64": dict(kind="soup", res=64, ntri=1500, W=96, H=64),
}


def _scene(name):
    sc = SceneSpec(**SCENES[name])
    if SCENES[name]["kind"] == "soup":
        sc.lights = [((0.3, 1.2, 0.5), (1.0, 1.0, 1.0), 2.0, 0)]
    sc.cone_set = L.CONES_REF9  # the reference's trace_diffuse
    return sc, np.load(GOLD / f"glref_{name}.npz")


def _bytes(a):
    return np.ascontiguousarray(a).view(np.uint8).reshape(-1, 4).astype(int)


def _split_mips(flat, res):
    out, off = {}, 0
    for l in range(1, 7):
        s = res >> l
        for d in range(6):
            out[(l, d)] = flat[off:off + s ** 3].reshape(s, s, s)
            off += s ** 3
    return out


@pytest.mark.parametrize("name", list(SCENES))
def test_voxelize_against_reference_shader(gpu_available, name):
    sc, g = _scene(name)
    print(f"voxelize vs llvmpipe: {name}", flush=True)
    with Context(32, 32) as ctx:
        sc.build(ctx, L.RGBA8, L.AVG_REFERENCE)
        ctx.voxelize()
        got = ctx.download_level(0)
    ref = g["level0"]
    assert np.array_equal(got != 0, ref != 0), "occupancy / coordinates differ from the reference shader"
    occ = ref != 0
    same = (got == ref)[occ].mean()
    worst = np.abs(_bytes(got[occ]) - _bytes(ref[occ])).max()
    print(f"  bit-identical voxels {same:.4f}, max byte difference {worst}")
    assert same > 0.99 and worst <= 2


@pytest.mark.parametrize("name", list(SCENES))
def test_mip_against_reference_shader(gpu_available, name):
    sc, g = _scene(name)
    print(f"mip vs llvmpipe: {name}", flush=True)
    ref = _split_mips(g["mips"], sc.res)
    with Context(32, 32) as ctx:
        ctx.set_grid(sc.res, 7, L.RGBA8)
        ctx.upload_level(0, 0, g["level0"])
        ctx.mip()
        worst, ndiff, n = 0, 0, 0
        for (l, d), r in ref.items():
            e = np.abs(_bytes(ctx.download_level(l, d)) - _bytes(r))
            worst = max(worst, int(e.max()))
            ndiff += int((e > 0).any(1).sum())
            n += e.shape[0]
    print(f"  max LSB difference {worst}, texels differing {ndiff}/{n}")
    assert worst <= 1 and ndiff < 0.02 * n


@pytest.mark.parametrize("name", list(SCENES))
def test_trace_against_reference_shader(gpu_available, name):
    """voxel_cone_tracing.frag (indirect light) on the reference's own volume
    and G-buffer vs the GPU cone marcher on the same inputs."""
    sc, g = _scene(name)
    print(f"trace vs llvmpipe: {name}", flush=True)
    ref_mips = _split_mips(g["mips"], sc.res)
    with Context(sc.W, sc.H) as ctx:
        sc.build(ctx, L.RGBA8, L.AVG_REFERENCE)
        ctx.upload_level(0, 0, g["level0"])
        for (l, d), a in ref_mips.items():
            ctx.upload_level(l, d, a)
        ctx.upload_gbuffer(g["gpos"], g["gnrm"])
        ctx.trace()
        img = ctx.download_image()
        ctx.gbuffer()
        gp, _ = ctx.download_gbuffer()
    cov = g["gpos"][..., 3] != 0
    e = img[cov][:, :3] - g["image"][cov][:, :3]
    psnr = 10 * np.log10(1.0 / np.mean(e.astype(np.float64) ** 2))
    print(f"  max-abs {np.abs(e).max():.3e}, PSNR {psnr:.1f} dB")
    assert psnr > 60.0 and np.abs(e).max() < 5e-3
    # the G-buffer raster covers the same pixels as GL's
    assert ((gp[..., 3] != 0) == cov).mean() > 0.995


def test_full_compose_against_reference_shader(gpu_available):
    """voxel_cone_tracing.frag with every phase on (direct light, shadow cones,
    diffuse and specular cones, emission, transmissive override) under
    llvmpipe vs the GPU marcher on the same volume and G-buffer."""
    g = np.load(GOLD / "glref_cornell64_full.npz")
    sc = SceneSpec("cornell", res=64, W=128, H=96, glass=True, phases=(True, True, True, True))
    sc.cone_set = L.CONES_REF9
    print("full compose vs llvmpipe: cornell64 glass", flush=True)
    ref_mips = _split_mips(g["mips"], sc.res)
    with Context(sc.W, sc.H) as ctx:
        sc.build(ctx, L.RGBA8, L.AVG_REFERENCE)
        ctx.upload_level(0, 0, g["level0"])
        for (l, d), a in ref_mips.items():
            ctx.upload_level(l, d, a)
        ctx.upload_gbuffer(g["gpos"], g["gnrm"])
        ctx.trace()
        img = ctx.download_image()
    cov = g["gpos"][..., 3] != 0
    e = img[cov][:, :3] - g["image"][cov][:, :3]
    psnr = 10 * np.log10(1.0 / np.mean(e.astype(np.float64) ** 2))
    print(f"  max-abs {np.abs(e).max():.3e}, PSNR {psnr:.1f} dB")
    assert psnr > 60.0 and np.abs(e).max() < 5e-3
