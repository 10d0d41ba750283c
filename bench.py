"""Benchmark: ms/frame of voxelize + mip + trace at 256^3, 2560x1440.

One step = one frame of the soup256 workload (1M-triangle soup, 256^3 grid
with 7 levels, FP16 radiance storage with FP32 accumulation, 2560x1440,
9 diffuse cones + 1 specular cone, indirect light only): voxelize, anisotropic
mip, G-buffer, cone trace, all on the GPU through libvct.so.  Geometry is
uploaded once before timing (resident in HBM).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload soup256]
                    [--no-cpu-baseline]

Prints one JSON line (see the keys below).  --gpus above 1 exits: this build
runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(wl, dtype, sample_tri=20000, sample_pix=1500):
    """The oracle (C++ port of the reference shaders) on a bounded sample of
    the same frame, scaled to ms/frame: voxelize on `sample_tri` triangles,
    the full mip chain, and the trace of `sample_pix` pixels."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle
    from voxel_cone_tracing_amd import _lib as L
    from voxel_cone_tracing_amd.scene import look_at, perspective, soup_materials_fast, identity
    from helpers import material_table

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    pos, nrm, ar = L.soup(wl.seed, sample_tri)
    mats = material_table(soup_materials_fast(ar))
    tm = np.arange(1, sample_tri + 1, dtype=np.uint32)
    t0 = time.perf_counter()
    v = oracle.voxelize(wl.res, wl.cube_size, identity(), pos, nrm, tm, mats, wl.lights, want=("q32",))
    t_vox = (time.perf_counter() - t0) * (wl.ntri / sample_tri)
    l0 = oracle.store_round(v["q32"], dtype)
    t0 = time.perf_counter()
    mips = oracle.mip_f32(wl.res, wl.levels, dtype, l0)
    t_mip = time.perf_counter() - t0
    # trace: G-buffer of the sample scene at a reduced size, pixels sampled from it
    W, H = 320, 180
    view = look_at(wl.eye, wl.target, (0, 1, 0))
    proj = perspective(wl.fovy, W / H, 0.1, 100.0)
    p32, n32, _, _, _ = oracle.gbuffer(W, H, view, proj, [(identity(), pos, nrm, tm)])
    cov = np.flatnonzero(p32[..., 3].reshape(-1) != 0)
    rng = np.random.default_rng(0)
    pix = rng.choice(cov, size=min(sample_pix, cov.size), replace=False).astype(np.int64)
    cam = view.reshape(16)[12:15]
    t0 = time.perf_counter()
    oracle.trace(np.float32, wl.res, wl.levels, l0, mips, p32, n32, mats, wl.lights, cam, wl.cube_size,
                 cone_set=wl.cone_set, pixels=pix)
    t_trace = (time.perf_counter() - t0) * (wl.width * wl.height / max(1, pix.size))
    total = t_vox + t_mip + t_trace
    return {
        "value": round(total * 1000.0, 1),
        "unit": "ms/frame",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle FP32 port: voxelize {sample_tri} of {wl.ntri} triangles scaled x{wl.ntri // sample_tri}, "
                   f"full {wl.levels}-level mip, trace {pix.size} sampled pixels scaled to {wl.width}x{wl.height}"),
        "parts_ms": {"voxelize": round(t_vox * 1e3, 1), "mip": round(t_mip * 1e3, 1), "trace": round(t_trace * 1e3, 1)},
    }


def mip_bytes(res, levels, bpe):
    """Compulsory HBM bytes of the fused mip: read level 0 once, write every
    mip texel of the six maps once."""
    rd = res ** 3 * bpe
    wr = sum(6 * (res >> l) ** 3 * bpe for l in range(1, levels))
    return rd + wr


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="soup256")
    ap.add_argument("--dtype", default="fp16", choices=["rgba8", "fp16", "bf16", "fp32"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (see DESIGN.md, sharding)",
              file=sys.stderr)
        sys.exit(2)

    from voxel_cone_tracing_amd import _lib as L
    from voxel_cone_tracing_amd.context import Context
    from voxel_cone_tracing_amd.scene import WORKLOADS, Scene

    dt = {"rgba8": L.RGBA8, "fp16": L.FP16, "bf16": L.BF16, "fp32": L.FP32}[args.dtype]
    wl = WORKLOADS[args.workload]
    log(f"bench: workload {wl.name} ({wl.ntri} tris, {wl.res}^3, {wl.width}x{wl.height}), dtype {args.dtype}")
    ctx = Context(wl.width, wl.height)
    t0 = time.perf_counter()
    sc = Scene(ctx, wl, dt)
    log(f"bench: scene uploaded in {time.perf_counter() - t0:.1f} s")

    def frame():
        sc.queue()
        ctx.render()

    for _ in range(args.warmup):
        frame()
    ctx.sync()
    log("bench: warmup done")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        frame()
    ctx.sync()
    elapsed = time.perf_counter() - t0
    ms = elapsed * 1000.0 / args.steps
    # per-pass HIP event times (recorded on the context stream), separate
    # frames so the timed loop above has no host synchronisation inside
    passes = np.zeros(5)
    for _ in range(args.steps):
        frame()
        passes += ctx.pass_times()
    passes /= args.steps
    log(f"bench: {ms:.3f} ms/frame; passes voxelize {passes[0]:.3f} mip {passes[1]:.3f} "
        f"gbuffer {passes[2]:.3f} trace {passes[3]:.3f} ms")
    nfrag, ncand = ctx.voxelize_stats()
    # one more trace with a step counter (not timed): sample count of the march
    frame()
    steps = ctx.trace_steps()
    samples = 3 * steps  # sample_voxel = 3 textureLod per step

    bpe = 8 if dt in (L.FP16, L.BF16) else (4 if dt == L.RGBA8 else 16)
    mb = mip_bytes(wl.res, wl.levels, bpe)
    mip_gbs = mb / (passes[1] * 1e-3) / 1e9 if passes[1] > 0 else 0.0
    dominant = int(np.argmax(passes[:4]))
    names = ["voxelize", "mip", "gbuffer", "trace"]
    roofline = {
        "kernel": "mip_fused_kernel (anisotropic mip chain)",
        "bound": "hbm",
        "achieved": round(mip_gbs, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(mip_gbs / HBM_PEAK_GBS, 4),
        "traffic": None,
        "note": f"dominant pass of the frame is {names[dominant]} ({passes[dominant]:.3f} ms)",
    }
    out = {
        "metric": "ms/frame voxelize+mip+trace @256^3,1440p",
        "value": round(ms, 4),
        "unit": "ms/frame",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{wl.name}: plain full-resolution form, {wl.ntri}-triangle soup, {wl.res}^3 grid "
                        f"({wl.levels} levels), {wl.width}x{wl.height}, 9 diffuse + 1 specular cone",
            "storage": args.dtype,
            "form": "plain",
        },
        "passes_ms": {n: round(float(passes[i]), 4) for i, n in enumerate(names)},
        "fragments": nfrag,
        "trace": {
            "cone_steps": steps,
            "samples": samples,
            "steps_per_pixel": round(steps / (wl.width * wl.height), 2),
            # each quadrilinear sample reads 16 texels (2 levels x 2x2x2)
            "effective_sample_GBps": round(samples * 16 * bpe / (passes[3] * 1e-3) / 1e9, 1) if passes[3] > 0 else 0,
        },
        "roofline": roofline,
    }
    if not args.no_cpu_baseline:
        log("bench: cpu baseline (oracle port, bounded sample)...")
        out["cpu_baseline"] = cpu_baseline(wl, dt)
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
