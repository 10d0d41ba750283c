"""Writes the llvmpipe reference fixtures used by the GPU tests.

Runs the reference's own shaders (read from /root/reference/shader at run
time, never copied) under Mesa llvmpipe via oracle/glref, single-threaded so
the output is reproducible, for small scenes built from seeds by
tests/helpers.SceneSpec, and stores the outputs only: level 0 and the mip
chain (RGBA8), the frame and the G-buffer (RGBA32F), with the llvmpipe pass
times; for the voxel views (view_voxel_dir < 7) the frame only.

    LP_NUM_THREADS=0 python tests/golden/make_glref_fixtures.py [scene ...]
"""
from __future__ import annotations

import os
import sys
from pathlib import Path

os.environ.setdefault("LP_NUM_THREADS", "0")
HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))
sys.path.insert(0, str(HERE.parent))

import numpy as np  # noqa: E402

from oracle import glref  # noqa: E402
from helpers import SceneSpec  # noqa: E402

SCENES = {
    "cornell64": dict(kind="cornell", res=64, W=128, H=96),
    "soup64": dict(kind="soup", res=64, ntri=1500, W=96, H=64),
    # every phase (direct light, shadow cones, diffuse and specular cones), a
    # transmissive glossy box and an emissive one: the full compose
    "cornell64_full": dict(kind="cornell", res=64, W=128, H=96, glass=True, phases=(True, True, True, True)),
    # the camera inside the box: walls that pass behind the eye are clipped at
    # the near plane (every phase on)
    "cornell64_inside": dict(kind="cornell", res=64, W=128, H=96, phases=(True, True, True, True),
                             eye=(0.12, 0.05, 0.2), target=(-0.25, -0.1, -0.5), fovy_deg=75.0),
    # a model matrix with rotation and non-uniform scale (voxelize.vert:24-30,
    # voxel_cone_tracing.vert:22-28)
    "soup64_xform": dict(kind="soup", res=64, ntri=1500, W=96, H=64, model="rot_scale"),
}


# The voxel view (view_voxel_dir < 7) of scenes above, image only (the volume
# is the base scene's): name -> (base scene, view_voxel_dir, view_voxel_lod)
VOXEL_VIEWS = {
    "vview_cornell64_d0": ("cornell64", 0, 0.0),
    "vview_cornell64_d3": ("cornell64", 3, 1.5),
    "vview_cornell64_inside_d1": ("cornell64_inside", 1, 0.5),
}


def reference_app_scene():
    """The scene of the reference application (main.cpp:85-121, 368-379):
    assets/CornellBox-Glossy.obj at identity and assets/suzanne.obj at
    translate(0, 1.1, -0.5) * scale(0.3) with the material main.cpp:89-98
    gives it through set_model_material; grid size 3, 128^3 (the Renderer's
    default), camera at (0, 0.9, 3) looking down -z with glm::perspective(45
    [radians], 4/3, 0.1, 100), the ceiling light; every phase on.  The
    suzanne transform is applied here (both sides then use identity), and
    the viewport is 400x300 (the application's 800x600 halved)."""
    from voxel_cone_tracing_amd import _lib as L
    from voxel_cone_tracing_amd.scene import identity, look_at, perspective
    from helpers import material_table

    assets = Path("/root/reference/assets")
    bp, bn, btm, bmats, _ = L.read_obj(assets / "CornellBox-Glossy.obj")
    sp, sn, stm, smats, _ = L.read_obj(assets / "suzanne.obj")
    # table: default (0), the box's MTL materials (1..), the suzanne material
    sz = L.make_material(diffuse=(0, 0, 0), specular=(1, 1, 1), emission=(0, 0, 0.25), transmittance=(1, 1, 1),
                         shininess=1000.0, ior=5.0, dissolve=0.1, illum=4)
    for k in range(3):
        sz.ambient[k] = 1.0
    mats = list(bmats) + [sz]
    box_tm = np.where(btm < 0, 0, btm).astype(np.uint32) + 1  # -1 takes the file's first material
    suz_tm = np.full(sp.shape[0], len(bmats) + 1, np.uint32)
    M = np.eye(4, dtype=np.float32)
    M[:3, :3] *= np.float32(0.3)
    M[:3, 3] = np.array([0, 1.1, -0.5], np.float32)
    v = sp.reshape(-1, 3).astype(np.float32)
    sp_w = (v @ M[:3, :3].T + M[:3, 3]).astype(np.float32).reshape(-1, 9)
    pos = np.concatenate([bp, sp_w]).astype(np.float32)
    nrm = np.concatenate([bn, sn]).astype(np.float32)
    tm = np.concatenate([box_tm, suz_tm]).astype(np.uint32)
    yaw = np.float32(np.radians(-90.0))
    front = np.array([np.cos(np.float32(0)) * np.cos(yaw), np.sin(np.float32(0)), np.cos(np.float32(0)) * np.sin(yaw)],
                     np.float32)
    front /= np.linalg.norm(front)
    eye = np.array([0, 0.9, 3], np.float32)
    view = look_at(eye, eye + front, (0, 1, 0))
    W, H = 400, 300
    proj = perspective(45.0, 800 / 600, 0.1, 100.0)
    lights = [((0.0, 1.4, 0.0), (1.0, 1.0, 1.0), 1.0, 0)]
    return dict(pos=pos, nrm=nrm, tri_mat=tm, mat_records=material_table(mats), M=identity(), view=view,
                proj=proj, W=W, H=H, res=128, cube_size=3.0, lights=lights, phases=(True, True, True, True),
                nbox=bp.shape[0], suzanne_local=sp, suzanne_M=M)


def soup_point_light(sc):
    sc.lights = [((0.3, 1.2, 0.5), (1.0, 1.0, 1.0), 2.0, 0)]


def main():
    only = sys.argv[1:]
    if not only or "reference_app" in only:
        sc = reference_app_scene()
        r = glref.frame(sc["res"], sc["cube_size"], sc["M"], sc["view"], sc["proj"], sc["W"], sc["H"], sc["pos"],
                        sc["nrm"], sc["tri_mat"], sc["mat_records"], sc["lights"], phases=sc["phases"])
        mips = np.concatenate([r["mips"][(l, d)].reshape(-1) for l in range(1, 7) for d in range(6)])
        out = HERE / "glref_reference_app.npz"
        np.savez_compressed(out, level0=r["level0"], mips=mips, image=r["image"], gpos=r["gpos"], gnrm=r["gnrm"],
                            times=r["times"], renderer=np.array(glref.renderer()), pos=sc["pos"], nrm=sc["nrm"],
                            tri_mat=sc["tri_mat"], mat_records=sc["mat_records"], view=sc["view"], proj=sc["proj"],
                            nbox=sc["nbox"], suzanne_local=sc["suzanne_local"], suzanne_M=sc["suzanne_M"])
        print(f"{out.name}: {out.stat().st_size / 1024:.0f} KiB, llvmpipe times {r['times']}")
    for name, kw in SCENES.items():
        if only and name not in only:
            continue
        sc = SceneSpec(**kw)
        if kw["kind"] == "soup":
            soup_point_light(sc)
        r = glref.frame(sc.res, sc.cube_size, sc.M, sc.view, sc.proj, sc.W, sc.H, sc.pos, sc.nrm, sc.tri_mat,
                        sc.mat_records, sc.lights, phases=sc.phases)
        mips = np.concatenate([r["mips"][(l, d)].reshape(-1) for l in range(1, 7) for d in range(6)])
        out = HERE / f"glref_{name}.npz"
        np.savez_compressed(out, level0=r["level0"], mips=mips, image=r["image"], gpos=r["gpos"], gnrm=r["gnrm"],
                            times=r["times"], renderer=np.array(glref.renderer()))
        print(f"{out.name}: {out.stat().st_size / 1024:.0f} KiB, llvmpipe times {r['times']}")
    for name, (base, vd, lod) in VOXEL_VIEWS.items():
        if only and name not in only:
            continue
        kw = SCENES[base]
        sc = SceneSpec(**kw)
        if kw["kind"] == "soup":
            soup_point_light(sc)
        r = glref.frame(sc.res, sc.cube_size, sc.M, sc.view, sc.proj, sc.W, sc.H, sc.pos, sc.nrm, sc.tri_mat,
                        sc.mat_records, sc.lights, phases=sc.phases, view_voxel_dir=vd, view_voxel_lod=lod)
        out = HERE / f"glref_{name}.npz"
        np.savez_compressed(out, image=r["image"], view_voxel_dir=vd, view_voxel_lod=lod, base=base,
                            renderer=np.array(glref.renderer()))
        print(f"{out.name}: {out.stat().st_size / 1024:.0f} KiB")


if __name__ == "__main__":
    main()
