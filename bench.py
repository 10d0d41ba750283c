"""Benchmark: ms/frame of voxelize + mip + trace at 256^3, 2560x1440, learned
quarter-resolution form vs plain full-resolution form, with PSNR against the
FP64 oracle.

One step = one frame of the soup256 workload (1M-triangle soup, 256^3 grid
with 7 levels, FP16 radiance storage with FP32 accumulation, 2560x1440,
9 diffuse cones + 1 specular cone, indirect light only), everything on the
GPU through libvct.so, geometry resident in HBM before timing:
  plain   : voxelize, anisotropic mip, G-buffer, full-resolution cone trace;
  learned : voxelize, mip, G-buffer, 640x360 cone trace, MLP upsampling to
            2560x1440, per-tile error estimate at probe pixels and
            full-resolution trace of the 32x32 tiles with the largest
            estimates until the frame holds 40 dB.
`value` is the ms/frame of the faster form that holds 40 dB against the
FP64 oracle (the learned form must also hold 40 dB against the plain frame);
both forms are reported, each with its PSNR against the FP64 oracle (voxelize, mip,
G-buffer and trace in double precision) over the covered pixels of a
512x512 centre crop.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload soup256]
                    [--weights file.npy] [--no-quality] [--no-cpu-baseline]

Prints one JSON line.  --gpus above 1 exits: the benchmark runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
WEIGHTS = ROOT / "voxel_cone_tracing_amd" / "weights" / "upsampler_soup256.npy"
TRAFFIC_PROFILE = ROOT / "profiles" / "hbm_traffic_soup256.md"
PASSES = ["voxelize", "mip", "gbuffer", "trace", "upsample", "fallback"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(wl, dtype):
    """The oracle (the C++ restatement of the reference's shaders, FP32) on the
    whole frame of the workload on the host's threads: voxelize of every
    triangle into the storage type, the full mip chain, the G-buffer at the
    full viewport and the trace of every covered pixel.  Nothing is sampled or
    scaled."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle
    from voxel_cone_tracing_amd import _lib as L
    from voxel_cone_tracing_amd.scene import look_at, perspective, soup_materials_fast, identity
    from helpers import material_table

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    pos, nrm, ar = L.soup(wl.seed, wl.ntri)
    mats = material_table(soup_materials_fast(ar))
    tm = np.arange(1, wl.ntri + 1, dtype=np.uint32)
    t0 = time.perf_counter()
    v = oracle.voxelize(wl.res, wl.cube_size, identity(), pos, nrm, tm, mats, wl.lights, want=("q32",))
    l0 = oracle.store_round(v["q32"], dtype)
    t_vox = time.perf_counter() - t0
    t0 = time.perf_counter()
    mips = oracle.mip_f32(wl.res, wl.levels, dtype, l0)
    t_mip = time.perf_counter() - t0
    view = look_at(wl.eye, wl.target, (0, 1, 0))
    proj = perspective(wl.fovy, wl.width / wl.height, 0.1, 100.0)
    t0 = time.perf_counter()
    p32, n32, _, _, _ = oracle.gbuffer(wl.width, wl.height, view, proj, [(identity(), pos, nrm, tm)])
    t_gb = time.perf_counter() - t0
    pix = np.flatnonzero(p32[..., 3].reshape(-1) != 0).astype(np.int64)
    t0 = time.perf_counter()
    oracle.trace(np.float32, wl.res, wl.levels, l0, mips, p32, n32, mats, wl.lights, view.reshape(16)[12:15],
                 wl.cube_size, cone_set=wl.cone_set, pixels=pix)
    t_trace = time.perf_counter() - t0
    total = t_vox + t_mip + t_gb + t_trace
    return {
        "value": round(total * 1000.0, 1),
        "unit": "ms/frame",
        "cores": threads,
        "kind": "port",
        "sample": (f"the whole frame, nothing scaled: oracle FP32 port, voxelize of all {wl.ntri} triangles, "
                   f"{wl.levels}-level mip, {wl.width}x{wl.height} G-buffer, trace of all {pix.size} covered "
                   f"pixels"),
        "parts_ms": {"voxelize": round(t_vox * 1e3, 1), "mip": round(t_mip * 1e3, 1), "gbuffer": round(t_gb * 1e3, 1),
                     "trace": round(t_trace * 1e3, 1)},
    }


def crop_window(wl, size=512):
    x0 = (wl.width // 2 - size // 2) // 32 * 32
    y0 = (wl.height // 2 - size // 2) // 32 * 32
    return x0, y0, size


def fp64_reference(sc, wl, crop):
    """FP64 oracle image (indirect light) of the workload frame on the crop."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle
    from voxel_cone_tracing_amd.scene import identity, soup_materials_fast
    from helpers import material_table

    # the oracle's table holds the renderer's default material at 0, the soup's records from 1
    mats = material_table(soup_materials_fast(sc.albedo_rough))
    tm = (sc.tri_mat - sc.tri_mat.min() + 1).astype(np.uint32)
    t0 = time.perf_counter()
    v = oracle.voxelize(wl.res, wl.cube_size, identity(), sc.pos, sc.nrm, tm, mats, wl.lights, want=("f64",))
    mips = oracle.mip_f64(wl.res, wl.levels, v["f64"])
    _, _, p64, n64, _ = oracle.gbuffer(wl.width, wl.height, sc.view, sc.proj, [(identity(), sc.pos, sc.nrm, tm)])
    x0, y0, s = crop
    yy, xx = np.mgrid[y0:y0 + s, x0:x0 + s]
    pix = (yy * wl.width + xx).reshape(-1).astype(np.int64)
    out, _ = oracle.trace(np.float64, wl.res, wl.levels, v["f64"], mips, p64, n64, mats, wl.lights,
                          sc.view.reshape(16)[12:15], wl.cube_size, cone_set=wl.cone_set, pixels=pix)
    cov = (p64[..., 3].reshape(-1)[pix] != 0).reshape(s, s)
    log(f"bench: FP64 oracle of the {s}x{s} crop in {time.perf_counter() - t0:.1f} s")
    return out.reshape(s, s, 4), cov


def psnr(a, b, cov):
    e = (a[..., :3].astype(np.float64) - b[..., :3])[cov]
    mse = float(np.mean(e * e))
    return (10.0 * math.log10(1.0 / mse) if mse > 0 else 999.0), float(np.abs(e).max())


def algorithmic(name, wl, bpe, steps, fb_tiles):
    """Algorithmic work of one pass: (amount, unit, bound, definition)."""
    W, H = wl.width, wl.height
    if name == "voxelize":
        return 84 * wl.ntri + wl.res ** 3 * bpe, "B", "hbm", "84 B per triangle + level 0 written once"
    if name == "mip":
        rd = wl.res ** 3 * bpe + sum(6 * (wl.res >> l) ** 3 * bpe for l in range(1, wl.levels - 1))
        wr = sum(6 * (wl.res >> l) ** 3 * bpe for l in range(1, wl.levels))
        if bpe == 16:
            return rd + wr, "B", "hbm", "each source level read once, every mip texel written once"
        vol = wl.res ** 3 * bpe + wr  # the x-paired copy: the whole volume written twice over
        return rd + wr + 2 * vol, "B", "hbm", ("each source level read once, every mip texel written once, and the "
                                                "x-paired copy (2x the volume) written")
    if name == "gbuffer":
        return 76 * wl.ntri + 32 * W * H, "B", "hbm", "76 B per triangle read + 32 B per pixel written"
    if name == "trace":
        return 3 * steps * 16 * bpe, "B", "vl1d", ("3 quadrilinear samples per cone step x 16 texels x bpp "
                                                   "(served from L1 / L2)")
    if name == "upsample":
        return 12928 * W * H, "FLOP", "mfma", "12,928 FLOP per output pixel (34-64-64-3 MLP)"
    return 0, "B", "hbm", ""


def copy_calibration(nbytes: int) -> None:
    """Two device-to-device copies of nbytes (the runtime's copy kernel), so
    that a counter run over this program also measures a copy of known size:
    its FETCH_SIZE / WRITE_SIZE per dispatch give the counters' scale."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) or hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, C.c_size_t(nbytes))
    for _ in range(2):
        if hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3):  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"bench: calibration copies of {nbytes} bytes done")


def sequential_passes(ctx, sc, steps):
    """Per-pass HIP event times of plain frames whose passes run one after the
    other on one stream (a frame overlaps the G-buffer with voxelize + mip, so
    its own pass times overlap)."""
    passes = np.zeros(6)
    for _ in range(steps):
        sc.queue()
        ctx.voxelize()
        ctx.mip()
        ctx.gbuffer()
        ctx.trace()
        passes += ctx.pass_times()
    ctx.clear_queues()
    return passes / steps


def run_form(ctx, sc, learned, steps, warmup):
    def frame():
        sc.queue()
        if learned:
            ctx.render_learned()
        else:
            ctx.render()

    # the timed frames record no pass-timing events (each is a packet the
    # stream processes between kernels); the per-pass HIP event times come
    # from the same frames run again with timing on, one synchronisation each
    ctx.set_pass_timing(0)
    for _ in range(warmup):
        frame()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        frame()
    ctx.sync()
    ms = (time.perf_counter() - t0) * 1000.0 / steps
    ctx.set_pass_timing(0x3F)
    passes = np.zeros(6)
    for _ in range(steps):
        frame()
        passes += ctx.pass_times()
    passes /= steps
    return ms, passes


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="soup256")
    ap.add_argument("--dtype", default="fp16", choices=["rgba8", "fp16", "bf16", "fp32"])
    ap.add_argument("--weights", default=str(WEIGHTS))
    ap.add_argument("--fallback-db", type=float, default=40.0)
    ap.add_argument("--no-quality", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="run two device copies of BYTES first (scale of the HBM byte counters)")
    ap.add_argument("--plain-only", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: the voxel cone tracing core has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)

    from voxel_cone_tracing_amd import _lib as L
    from voxel_cone_tracing_amd.context import Context
    from voxel_cone_tracing_amd.scene import WORKLOADS, Scene

    dt = {"rgba8": L.RGBA8, "fp16": L.FP16, "bf16": L.BF16, "fp32": L.FP32}[args.dtype]
    bpe = 8 if dt in (L.FP16, L.BF16) else (4 if dt == L.RGBA8 else 16)
    wl = WORKLOADS[args.workload]
    log(f"bench: workload {wl.name} ({wl.ntri} tris, {wl.res}^3, {wl.width}x{wl.height}), dtype {args.dtype}")
    ctx = Context(wl.width, wl.height)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
    t0 = time.perf_counter()
    sc = Scene(ctx, wl, dt)
    log(f"bench: scene uploaded in {time.perf_counter() - t0:.1f} s")

    # ---- plain form
    plain_ms, plain_passes = run_form(ctx, sc, False, args.steps, args.warmup)
    log(f"bench: plain {plain_ms:.3f} ms/frame; passes " +
        " ".join(f"{n} {plain_passes[i]:.3f}" for i, n in enumerate(PASSES[:4])))
    sc.queue()
    ctx.render()
    plain_img = ctx.download_image()
    gp, _ = ctx.download_gbuffer()
    cov_full = gp[..., 3] != 0
    steps = ctx.trace_steps()

    # ---- learned form
    have_weights = Path(args.weights).exists() and not args.plain_only
    learned = None
    if have_weights:
        ctx.mlp_set_params(np.load(args.weights))
        ctx.set_fallback(args.fallback_db)
        l_ms, l_passes = run_form(ctx, sc, True, args.steps, args.warmup)
        sc.queue()
        ctx.render_learned()
        l_img = ctx.download_image()
        nt, nf = ctx.fallback_stats()
        traced = ctx.traced_pixels()
        stats = ctx.upsample_stats()
        ctx.set_fallback(0.0)
        sc.queue()
        ctx.render_learned()
        raw_img = ctx.download_image()
        ctx.set_fallback(args.fallback_db)
        learned = dict(ms=l_ms, passes=l_passes, img=l_img, raw=raw_img, tiles=nt, fb=nf, traced=traced, stats=stats)
        log(f"bench: learned {l_ms:.3f} ms/frame; passes " +
            " ".join(f"{n} {l_passes[i]:.3f}" for i, n in enumerate(PASSES)) + f"; fallback tiles {nf}/{nt}")
    elif args.plain_only:
        log("bench: plain form only (--plain-only)")
    else:
        log(f"bench: no upsampler parameters at {args.weights}: plain form only")

    # ---- quality
    quality = {}
    if learned is not None:
        quality["learned_vs_plain_full_frame_db"] = round(psnr(learned["img"], plain_img, cov_full)[0], 2)
        quality["learned_no_fallback_vs_plain_full_frame_db"] = round(psnr(learned["raw"], plain_img, cov_full)[0], 2)
    if not args.no_quality and wl.name.startswith("soup"):
        crop = crop_window(wl, 256 if wl.res >= 512 else 512)  # 256x256 crop at 512^3
        ref, cov = fp64_reference(sc, wl, crop)
        x0, y0, s = crop
        sl = (slice(y0, y0 + s), slice(x0, x0 + s))
        p, m = psnr(plain_img[sl], ref, cov)
        quality["crop"] = f"{s}x{s} at ({x0},{y0}), {int(cov.sum())} covered pixels"
        quality["plain_vs_fp64_db"] = round(p, 2)
        quality["plain_vs_fp64_max_abs"] = float(f"{m:.3e}")
        if learned is not None:
            p, m = psnr(learned["img"][sl], ref, cov)
            quality["learned_vs_fp64_db"] = round(p, 2)
            quality["learned_vs_fp64_max_abs"] = float(f"{m:.3e}")
            p, _ = psnr(learned["raw"][sl], ref, cov)
            quality["learned_no_fallback_vs_fp64_db"] = round(p, 2)

    # ---- roofline of the cone marcher, the kernel that dominates both forms: the
    # plain frame's trace pass is exactly one trace_frame_kernel launch, timed by
    # HIP events on the context stream around it (the learned form runs the same
    # marcher over the quarter-resolution frame and over tile lists).  Its texel
    # gathers are served from L1/L2, so its ceiling is the rate at which the
    # vector memory path delivers its 16-byte loads from L1, measured here by a
    # calibration kernel of the same load form (vct_trace_load_peak).
    amount, unit, bound, definition = algorithmic("trace", wl, bpe, steps, 0)
    t_launch = float(plain_passes[3]) * 1e-3
    achieved = amount / t_launch / 1e9
    load_peak = ctx.load_peak_gbps()
    log(f"bench: L1-hit 16-byte load ceiling {load_peak:.0f} GB/s")
    # up to 9 diffuse cones without shadow lights: the 128-pixel (16x8) block shape
    kname = f"trace_frame_kernel<{dt},Shape<128,2>>"
    grid = (wl.width + 15) // 16 * 256 * ((wl.height + 7) // 8)
    traffic = l2_hit = None
    from voxel_cone_tracing_amd.profile_summary import read_summary, read_traffic, summary_is_current

    if wl.name == "soup256" and args.dtype == "fp16":
        # counters of a profile taken with the kernel sources as they are now, else none
        traffic = read_traffic(TRAFFIC_PROFILE, kname, grid)
        l2_hit = read_summary(TRAFFIC_PROFILE, "l2_hit_rate", kname, grid)
        if not summary_is_current(TRAFFIC_PROFILE):
            log(f"bench: {TRAFFIC_PROFILE.name} was taken with other kernel sources: no traffic figure")
    roofline = {
        "kernel": "trace_frame_kernel", "pass": "trace (plain frame, one launch)", "bound": "vl1d",
        "achieved": round(achieved, 1), "peak": round(load_peak, 1), "unit": "GB/s",
        "frac": round(achieved / load_peak, 4), "traffic": traffic, "l2_hit_rate": l2_hit,
        "hbm_GBps": round(traffic / t_launch / 1e9, 1) if traffic else None,
        "algorithmic": definition, "launch_ms": round(float(plain_passes[3]), 4),
        "note": ("bound vl1d: the texel gathers hit L1/L2; peak = 16-byte loads hitting L1, measured live by a "
                 "calibration kernel of the same load form (no gather pattern of these loads is delivered "
                 "faster); traffic = HBM bytes per launch from rocprofv3 FETCH_SIZE/WRITE_SIZE (scaled by a "
                 "copy-kernel calibration) of a profile stamped with the same kernel sources, else null"),
    }
    # HBM-bound passes against the HBM peak (algorithmic bytes / pass time), timed
    # in frames whose passes do not overlap
    seq = sequential_passes(ctx, sc, max(3, args.steps // 4))
    hbm_passes = {}
    for i, n in ((0, "voxelize"), (1, "mip")):
        amt, _, _, dfn = algorithmic(n, wl, bpe, steps, 0)
        if seq[i] > 0:
            gbs = amt / (float(seq[i]) * 1e-3) / 1e9
            hbm_passes[n] = {"achieved_GBps": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4),
                             "bytes": int(amt), "algorithmic": dfn, "pass_ms": round(float(seq[i]), 4)}
    # value: the faster form that holds the 40 dB bound (the plain form holds it
    # by construction: it is the full-resolution trace)
    learned_ok = learned is not None and quality.get("learned_vs_plain_full_frame_db", 0.0) >= 40.0 and \
        quality.get("learned_vs_fp64_db", 99.0) >= 40.0
    use_learned = learned_ok and learned["ms"] < plain_ms
    value = learned["ms"] if use_learned else plain_ms
    out = {
        "metric": "ms/frame voxelize+mip+trace @256^3,1440p; learned vs plain at >=40 dB PSNR",
        "value": round(value, 4),
        "unit": "ms/frame",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(value, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{wl.name}: {wl.ntri}-triangle soup, {wl.res}^3 grid ({wl.levels} levels, {args.dtype} "
                        f"storage), {wl.width}x{wl.height}, 9 diffuse + 1 specular cone, indirect light",
            "form": "learned (640x360 trace + MLP upsampling + per-tile fallback)" if use_learned else
                    "plain (full-resolution trace; the faster form at >= 40 dB)",
        },
        "plain_ms_per_frame": round(plain_ms, 4),
        "plain_passes_ms": {n: round(float(plain_passes[i]), 4) for i, n in enumerate(PASSES[:4])},
        "plain_passes_sequential_ms": {n: round(float(seq[i]), 4) for i, n in enumerate(PASSES[:4])},
    }
    if learned is not None:
        out["learned_ms_per_frame"] = round(learned["ms"], 4)
        out["learned_passes_ms"] = {n: round(float(learned["passes"][i]), 4) for i, n in enumerate(PASSES)}
        out["speedup_learned_vs_plain"] = round(plain_ms / learned["ms"], 3)
        out["fallback_tiles"] = f"{learned['fb']}/{learned['tiles']}"
        out["traced_pixels"] = f"{learned['traced']}/{int(cov_full.sum())} covered"
        out["learned_stats"] = learned["stats"]
    out["quality"] = quality
    out["trace"] = {
        "cone_steps": steps,
        "samples": 3 * steps,
        "steps_per_pixel": round(steps / (wl.width * wl.height), 2),
        "effective_sample_GBps": round(3 * steps * 16 * bpe / (plain_passes[3] * 1e-3) / 1e9, 1)
        if plain_passes[3] > 0 else 0,
    }
    out["roofline"] = roofline
    out["hbm_passes"] = hbm_passes
    if not args.no_cpu_baseline:
        log("bench: cpu baseline (oracle port, the whole frame)...")
        out["cpu_baseline"] = cpu_baseline(wl, dt)
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
