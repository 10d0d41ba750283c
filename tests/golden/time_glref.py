"""Times the reference's shaders under Mesa llvmpipe (the CPU path of the
reference) on the Cornell box and on soups at the soup256 size.

    LP_NUM_THREADS=8 python tests/golden/time_glref.py [--soup-tris 20000 1000000] [--buckets 64]

Needs the reference tree and Mesa's swrast driver (this container only).
Prints one line per frame: pass times [voxelize, mip, trace] in seconds.

The reference binds a material per draw call (src/renderer.cpp:248-255), so a
soup with per-triangle albedo costs one draw per triangle under its
unmodified shaders.  --buckets K gives the soup K materials instead (the
first K triangles' albedo and roughness, triangle t takes material t mod K)
and submits the triangles grouped by material: K draws per pass, the form a
scene of K materials has in the reference.
"""
from __future__ import annotations

import argparse
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

import numpy as np  # noqa: E402

from oracle import glref  # noqa: E402
from helpers import SceneSpec, material_table  # noqa: E402
from voxel_cone_tracing_amd import _lib as L  # noqa: E402
from voxel_cone_tracing_amd.scene import WORKLOADS, identity, look_at, perspective, soup_materials_fast  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--soup-tris", type=int, nargs="*", default=[20000])
    ap.add_argument("--frames", type=int, default=2)
    ap.add_argument("--buckets", type=int, default=0, help="K materials, one draw each (0: one per triangle)")
    args = ap.parse_args()
    if not glref.available():
        sys.exit("llvmpipe harness unavailable (needs the reference tree and Mesa swrast)")
    print(glref.renderer(), flush=True)
    sc = SceneSpec("cornell", res=64, W=512, H=512)
    for _ in range(args.frames + 1):
        out = glref.frame(64, sc.cube_size, sc.M, sc.view, sc.proj, 512, 512, sc.pos, sc.nrm, sc.tri_mat,
                          sc.mat_records, sc.lights, want=("image",))
        print("cornell64 512x512 [vox, mip, trace] s:", [round(float(t), 4) for t in out["times"]], flush=True)
    wl = WORKLOADS["soup256"]
    view = look_at(wl.eye, wl.target, (0, 1, 0))
    proj = perspective(wl.fovy, wl.width / wl.height, 0.1, 100.0)
    lights = [((0.4, 1.2, 0.6), (1.0, 1.0, 1.0), 1.0, 0)]  # the shaders have point lights only
    for n in args.soup_tris:
        pos, nrm, ar = L.soup(wl.seed, n)
        if args.buckets:
            k = args.buckets
            mats = material_table(soup_materials_fast(ar[:k]))
            order = np.argsort(np.arange(n) % k, kind="stable")  # grouped by material: one draw per bucket
            pos, nrm = np.ascontiguousarray(pos[order]), np.ascontiguousarray(nrm[order])
            tm = (order % k + 1).astype(np.uint32)
        else:
            mats = material_table(soup_materials_fast(ar))
            tm = np.arange(1, n + 1, dtype=np.uint32)
        for _ in range(args.frames):
            t0 = time.time()
            out = glref.frame(wl.res, wl.cube_size, identity(), view, proj, wl.width, wl.height, pos, nrm, tm, mats,
                              lights, want=("image",))
            print(f"soup {n} tris ({args.buckets or n} materials) {wl.res}^3 {wl.width}x{wl.height} "
                  f"[vox, mip, trace] s:",
                  [round(float(t), 4) for t in out["times"]], f"wall {time.time() - t0:.1f}", flush=True)


if __name__ == "__main__":
    main()
