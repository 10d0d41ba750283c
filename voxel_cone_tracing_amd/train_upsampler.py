"""Train the quarter-resolution upsampler on the GPU.

Training data are full-resolution frames that the plain cone tracer renders
of seeded soups (the soup256 distributions: 1M triangles, 256^3 grid, FP16
radiance, 2560x1440, 9 diffuse + 1 specular cone) from seeded random camera
poses.  Every frame contributes `--samples` random pixels (features, albedo,
albedo-demodulated target) to a device-resident training set; the network is
then fitted with the library's own BF16 forward/backward kernels and Adam.

Held-out evaluation: the soup256 benchmark scene (a scene seed that is not
trained on) from its benchmark camera and from random poses with pose seeds
>= 10000; PSNR over covered pixels against the plain full-resolution GPU
frame, for the learned frame without the tile fallback, for the same frame
with a zero network (the surface-guided interpolation alone), for bilinear
upsampling of the same quarter-resolution trace; the pixels traced at full
resolution, and the share of 32x32 tiles whose PSNR is below 40 dB.

    python -m voxel_cone_tracing_amd.train_upsampler --frames 4096 --out weights.npy --record train.json
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import math
import sys
import time

import numpy as np

from . import _lib as L
from .context import Context
from .scene import WORKLOADS, Scene, look_at, perspective


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def random_pose(rng):
    """Eye on a shell around the soup, elevation within +-50 degrees, looking
    near the centre."""
    az = rng.uniform(0.0, 2.0 * math.pi)
    el = rng.uniform(-50.0, 50.0) * math.pi / 180.0
    r = rng.uniform(2.4, 4.0)
    eye = (r * math.cos(el) * math.sin(az), r * math.sin(el), r * math.cos(el) * math.cos(az))
    target = tuple(rng.uniform(-0.3, 0.3, 3))
    return eye, target


def bilinear_upsample(low: np.ndarray, W: int, H: int) -> np.ndarray:
    """Bilinear interpolation of the quarter-resolution buffer at every
    full-resolution pixel (low-res pixel (i, j) sits at full-res (4i+1, 4j+1))."""
    hl, wl = low.shape[:2]
    u = (np.arange(W) - 1) / 4.0
    v = (np.arange(H) - 1) / 4.0
    i0 = np.clip(np.floor(u).astype(int), 0, wl - 1)
    j0 = np.clip(np.floor(v).astype(int), 0, hl - 1)
    i1 = np.clip(i0 + 1, 0, wl - 1)
    j1 = np.clip(j0 + 1, 0, hl - 1)
    au = np.clip(u - np.floor(u), 0, 1)[None, :, None]
    av = np.clip(v - np.floor(v), 0, 1)[:, None, None]
    a = low[j0][:, i0, :3]
    b = low[j0][:, i1, :3]
    c = low[j1][:, i0, :3]
    d = low[j1][:, i1, :3]
    return (1 - au) * (1 - av) * a + au * (1 - av) * b + (1 - au) * av * c + au * av * d


def psnr_covered(a, b, cov):
    e = (a[..., :3] - b[..., :3])[cov].astype(np.float64)
    mse = float(np.mean(e * e)) if e.size else 0.0
    return 10.0 * math.log10(1.0 / mse) if mse > 0 else float("inf"), mse


def tile_fail_share(a, b, cov, tile=32, psnr_db=40.0):
    H, W = cov.shape
    bad = tot = 0
    lim = 10.0 ** (-psnr_db / 10.0)
    for y in range(0, H, tile):
        for x in range(0, W, tile):
            c = cov[y:y + tile, x:x + tile]
            if not c.any():
                continue
            e = (a[y:y + tile, x:x + tile, :3] - b[y:y + tile, x:x + tile, :3])[c]
            tot += 1
            bad += float(np.mean(e.astype(np.float64) ** 2)) > lim
    return bad / max(tot, 1)


def albedo_image(ctx, sc) -> np.ndarray:
    """Per-pixel diffuse albedo from the G-buffer's material ids."""
    _, gn = ctx.download_gbuffer()
    mid = gn[..., 3].view(np.uint32).astype(np.int64)
    alb = np.zeros(gn.shape[:2] + (3,), np.float32)
    first = int(sc.tri_mat.min())
    idx = np.clip(mid - first, 0, sc.albedo_rough.shape[0] - 1)
    alb[:] = sc.albedo_rough[idx, :3]
    return alb


def evaluate(ctx, sc, wl, poses):
    rows = []
    for (eye, target) in poses:
        view = look_at(eye, target, (0, 1, 0))
        proj = perspective(wl.fovy, wl.width / wl.height, 0.1, 100.0)
        ctx.set_camera(view, proj)
        sc.queue()
        ctx.render()
        plain = ctx.download_image()
        gp, _ = ctx.download_gbuffer()
        cov = gp[..., 3] != 0
        ctx.set_fallback(0.0)
        sc.queue()
        ctx.render_learned()
        learned = ctx.download_image()
        traced = ctx.traced_pixels()
        low = ctx.download_lowres()
        alb = albedo_image(ctx, sc)
        bil = bilinear_upsample(low, wl.width, wl.height) * alb
        params = ctx.mlp_params()
        ctx.mlp_set_params(np.zeros_like(params))
        sc.queue()
        ctx.render_learned()
        guided = ctx.download_image()
        ctx.mlp_set_params(params)
        p_l, _ = psnr_covered(learned, plain, cov)
        p_g, _ = psnr_covered(guided, plain, cov)
        p_b, _ = psnr_covered(bil, plain, cov)
        fail = tile_fail_share(learned, plain, cov)
        ctx.set_fallback(40.0)
        sc.queue()
        ctx.render_learned()
        fb = ctx.download_image()
        p_f, _ = psnr_covered(fb, plain, cov)
        nt, nf = ctx.fallback_stats()
        rows.append({"eye": [round(float(x), 4) for x in eye], "psnr_learned_db": round(p_l, 2),
                     "psnr_guided_interpolation_db": round(p_g, 2), "psnr_bilinear_db": round(p_b, 2),
                     "traced_pixels": traced, "covered_pixels": int(cov.sum()),
                     "tiles_below_40db": round(fail, 4), "psnr_with_fallback_db": round(p_f, 2),
                     "fallback_tiles": nf, "tiles": nt})
        log(f"  eval eye={rows[-1]['eye']}: learned {p_l:.2f} dB, guided {p_g:.2f} dB, bilinear {p_b:.2f} dB, "
            f"traced px {traced}/{int(cov.sum())}, tiles<40dB {fail:.3f}, with fallback {p_f:.2f} dB "
            f"({nf}/{nt} tiles)")
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="soup256")
    ap.add_argument("--scenes", type=int, nargs="+", default=[101, 102, 103, 104])
    ap.add_argument("--frames", type=int, default=4096, help="training frames over all scenes")
    ap.add_argument("--samples", type=int, default=4096, help="pixels sampled per frame")
    ap.add_argument("--steps", type=int, default=16000)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--init", default=None, help="start from these parameters (.npy)")
    ap.add_argument("--eval-poses", type=int, default=4)
    ap.add_argument("--out", required=True)
    ap.add_argument("--record", default=None)
    args = ap.parse_args()

    wl0 = WORKLOADS[args.workload]
    ctx = Context(wl0.width, wl0.height)
    t_start = time.perf_counter()
    if args.init:
        ctx.mlp_set_params(np.load(args.init))
    else:
        ctx.mlp_init(1)
    per_scene = max(1, args.frames // len(args.scenes))
    t0 = time.perf_counter()
    frame = 0
    for s_i, seed in enumerate(args.scenes):
        wl = dataclasses.replace(wl0, seed=seed)
        sc = Scene(ctx, wl, L.FP16)
        sc.queue()
        ctx.voxelize()
        ctx.mip()
        rng = np.random.default_rng(seed)
        for k in range(per_scene):
            eye, target = random_pose(rng)
            ctx.set_camera(look_at(eye, target, (0, 1, 0)), perspective(wl.fovy, wl.width / wl.height, 0.1, 100.0))
            ctx.gbuffer()
            ctx.upsample_collect(args.samples, seed * 1_000_003 + k)
            frame += 1
            if frame % 256 == 0:
                log(f"collect: {frame} frames, {ctx.upsample_count()} samples, {time.perf_counter() - t0:.1f} s")
        ctx.clear_queues()
    t_collect = time.perf_counter() - t0
    log(f"collected {ctx.upsample_count()} samples from {frame} frames in {t_collect:.1f} s")

    t0 = time.perf_counter()
    losses = []
    phases = 4
    for ph in range(phases):
        lr = args.lr * (0.5 ** ph)
        n = args.steps // phases
        chunk = 1000
        for c0 in range(0, n, chunk):
            loss = ctx.upsample_train(min(chunk, n - c0), args.batch, lr, seed=ph * 1_000_000 + c0)
            losses.append(loss)
            log(f"train: phase {ph} lr {lr:.2e} step {c0 + min(chunk, n - c0)}/{n} loss {loss:.3e}")
    t_train = time.perf_counter() - t0
    params = ctx.mlp_params()
    np.save(args.out, params)
    log(f"trained {args.steps} steps in {t_train:.1f} s; saved {args.out}")

    # held-out evaluation on the benchmark scene
    sc = Scene(ctx, wl0, L.FP16)
    rng = np.random.default_rng(10000)
    poses = [(wl0.eye, wl0.target)] + [random_pose(rng) for _ in range(args.eval_poses - 1)]
    rows = evaluate(ctx, sc, wl0, poses)
    rec = {
        "workload": wl0.name, "train_scenes": args.scenes, "frames": frame, "samples": int(ctx.upsample_count()),
        "samples_per_frame": args.samples, "steps": args.steps, "batch": args.batch, "lr": args.lr,
        "loss_trace": [round(float(x), 6) for x in losses], "collect_s": round(t_collect, 1),
        "train_s": round(t_train, 1), "total_s": round(time.perf_counter() - t_start, 1),
        "eval_scene_seed": wl0.seed, "eval": rows,
    }
    if args.record:
        with open(args.record, "w") as f:
            json.dump(rec, f, indent=1)
    print(json.dumps({k: v for k, v in rec.items() if k != "loss_trace"}))
    ctx.close()


if __name__ == "__main__":
    main()
