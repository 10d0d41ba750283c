"""Writes the llvmpipe reference fixtures used by the GPU tests.

Runs the reference's own shaders (read from /root/reference/shader at run
time, never copied) under Mesa llvmpipe via oracle/glref, single-threaded so
the output is reproducible, for two small scenes built from seeds by
tests/helpers.SceneSpec, and stores the outputs only: level 0 and the mip
chain (RGBA8), the indirect-light image and the G-buffer (RGBA32F), with
the llvmpipe pass times.

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
}


def soup_point_light(sc):
    sc.lights = [((0.3, 1.2, 0.5), (1.0, 1.0, 1.0), 2.0, 0)]


def main():
    only = sys.argv[1:]
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


if __name__ == "__main__":
    main()
