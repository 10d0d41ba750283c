"""Build recipe of libvct.so (the gfx950 HIP kernels + the C ABI of include/vct.h).

Every source under csrc/ is compiled by hipcc for gfx950 only, in parallel,
into csrc/_build/, then linked into voxel_cone_tracing_amd/libvct.so.
Kernels are built with packed-FP32 instruction selection switched off.

    python -m voxel_cone_tracing_amd.build        # build (always recompiles)
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT_DIR = CSRC / "_build"
LIB = PKG / "libvct.so"
DEMO = PKG / "vct_demo"
INCLUDE = PKG.parent / "include"

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Wno-unused-but-set-variable",
    f"-I{INCLUDE}",
    # no packed FP32 instructions (v_pk_*_f32) in any kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _sources() -> list[Path]:
    return sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))


def _compile(src: Path) -> Path:
    obj = OUT_DIR / (src.name + ".o")
    cmd = [HIPCC, *COMMON, "-x", "hip", "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> Path:
    """Compile every HIP/C++ source for gfx950 and link libvct.so."""
    if not Path(HIPCC).exists() and shutil.which("hipcc") is None:
        raise RuntimeError("hipcc not found")
    OUT_DIR.mkdir(parents=True, exist_ok=True)
    srcs = _sources()
    jobs = min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    # host programs on the Renderer API (vct_demo: the headless frame loop;
    # vct_reference_app: the reference application's scene), next to libvct.so
    for app in sorted((PKG / "apps").glob("*.cpp")):
        exe = PKG / app.stem
        cmd = ["g++", "-O2", "-std=c++17", f"-I{INCLUDE}", str(app), "-o", str(exe), f"-L{PKG}", "-lvct",
               "-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"{app.stem} build failed:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources")
    return LIB


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
