"""The C++ drop-in (include/vct_renderer.h) run on the GPU by host programs.

vct_reference_app sets up the reference application's scene the way
src/main.cpp:85-121 does - load_model of the Cornell box and of suzanne,
add_material / set_model_material, set_grid_size, the camera, the ceiling
light - and runs the body of its frame loop (main.cpp:300-387) through
Renderer, then writes Renderer::image().  The reference's asset files are not
part of this repository, so the test writes the triangulated geometry and the
materials the llvmpipe fixture holds as OBJ/MTL files and load_model reads
them (renderer.cpp:407-603).  The image is compared with the reference's own
frame under llvmpipe (tests/golden/glref_reference_app.npz).

vct_demo (the headless frame loop of the procedural Cornell box) is compared
with the same frame built through the C ABI from Python.
"""
from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np
import pytest

from voxel_cone_tracing_amd import _lib as L
from voxel_cone_tracing_amd.context import Context
from voxel_cone_tracing_amd.scene import look_at, perspective

pytestmark = pytest.mark.gpu
PKG = Path(L.__file__).resolve().parent
GOLD = Path(__file__).resolve().parent / "golden"


def _f(x) -> str:
    return f"{float(x):.9g}"


def write_mtl(path: Path, records: np.ndarray) -> list[str]:
    """material_data_t records (n, 32) as MTL materials m0..m{n-1}: the fields
    load_model reads (Ka Kd Ks Tf Ke Ns Ni d illum)."""
    names, lines = [], []
    for k, r in enumerate(records):
        names.append(f"m{k}")
        illum = int(np.array([r[22]], np.float32).view(np.int32)[0])
        lines += [f"newmtl m{k}", "Ka " + " ".join(_f(v) for v in r[0:3]), "Kd " + " ".join(_f(v) for v in r[4:7]),
                  "Ks " + " ".join(_f(v) for v in r[8:11]), "Tf " + " ".join(_f(v) for v in r[12:15]),
                  "Ke " + " ".join(_f(v) for v in r[16:19]), f"Ns {_f(r[19])}", f"Ni {_f(r[20])}",
                  f"d {_f(r[21])}", f"illum {illum}", ""]
    path.write_text("\n".join(lines))
    return names


def write_obj(path: Path, pos: np.ndarray, nrm: np.ndarray, tri_mat=None, mtl: str | None = None, names=None):
    """Triangles (9 floats each) with per-vertex normals; tri_mat indexes names."""
    p = pos.reshape(-1, 3)
    n = nrm.reshape(-1, 3)
    out = [f"mtllib {mtl}"] if mtl else []
    out += ["v " + " ".join(_f(c) for c in v) for v in p]
    out += ["vn " + " ".join(_f(c) for c in v) for v in n]
    cur = None
    for t in range(pos.reshape(-1, 9).shape[0]):
        if tri_mat is not None and tri_mat[t] != cur:
            cur = tri_mat[t]
            out.append(f"usemtl {names[cur]}")
        a, b, c = 3 * t + 1, 3 * t + 2, 3 * t + 3
        out.append(f"f {a}//{a} {b}//{b} {c}//{c}")
    path.write_text("\n".join(out) + "\n")


def psnr(a, b, cov):
    e = a[cov][:, :3].astype(np.float64) - b[cov][:, :3]
    return 10 * np.log10(1.0 / np.mean(e ** 2)), float(np.abs(e).max())


def test_reference_app_through_renderer(gpu_available, tmp_path):
    """main.cpp's scene and frame through Renderer (load_model included) vs the
    reference's frame under llvmpipe."""
    g = np.load(GOLD / "glref_reference_app.npz")
    nb = int(g["nbox"])
    box_tm = g["tri_mat"][:nb].astype(np.int64)
    k = g["mat_records"].shape[0] - 2  # the table: default, the box's k materials, suzanne's
    names = write_mtl(tmp_path / "box.mtl", g["mat_records"][1:k + 1])
    write_obj(tmp_path / "box.obj", g["pos"][:nb], g["nrm"][:nb], box_tm - 1, "box.mtl", names)
    write_obj(tmp_path / "suzanne.obj", g["suzanne_local"], g["nrm"][nb:])
    H, W = g["image"].shape[:2]
    out = tmp_path / "frame.f32"
    print(f"vct_reference_app {W}x{H}", flush=True)
    r = subprocess.run([str(PKG / "vct_reference_app"), str(tmp_path / "box.obj"), str(tmp_path / "suzanne.obj"),
                        str(W), str(H), str(out), "2"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    img = np.fromfile(out, np.float32).reshape(H, W, 4)
    cov = g["gpos"][..., 3] != 0
    p, m = psnr(img, g["image"], cov)
    print(f"  Renderer frame vs llvmpipe: {p:.1f} dB, max-abs {m:.3e}")
    assert np.isfinite(img).all()
    assert p > 50.0


def test_demo_frame_matches_the_c_abi(gpu_available, tmp_path):
    """vct_demo (Renderer, procedural Cornell box, 512x512, every phase) against
    the same frame set up through the C ABI."""
    out = tmp_path / "demo.f32"
    r = subprocess.run([str(PKG / "vct_demo"), "--frames", "2", "--raw", str(out)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0 and "checksum" in r.stdout
    W = H = 512
    demo = np.fromfile(out, np.float32).reshape(H, W, 4)
    pos, nrm, tm, mats = L.cornell(1)
    with Context(W, H) as ctx:
        ctx.set_grid(64, 7, L.RGBA8, L.AVG_REFERENCE)
        ctx.set_grid_size(0.55)
        base = ctx.add_materials(mats)
        model = ctx.model_create(pos, nrm, (tm + base).astype(np.uint32))
        ctx.set_camera(look_at((0.0, 0.0, 1.55), (0, 0, 0), (0, 1, 0)),
                       perspective(45.0 * 3.14159265 / 180.0, 1.0, 0.1, 100.0))
        ctx.set_phases(True, True, True, True)
        ctx.set_voxel_view(7, 0.0)
        ctx.queue_model(model)
        ctx.queue_light((0.0, 0.45, 0.0), (1, 1, 1), 1.0, L.LIGHT_POINT)
        ctx.render()
        img = ctx.download_image()
        gp, _ = ctx.download_gbuffer()
    cov = gp[..., 3] != 0
    p, m = psnr(demo, img, cov)
    print(f"  vct_demo vs the C ABI frame: {p:.1f} dB, max-abs {m:.3e}")
    assert cov.mean() > 0.5
    assert p > 80.0 and m < 1e-3
