"""GPU parity of the learned upsampler's matrix-core kernels (through the C ABI).

Forward outputs and the loss gradient of every parameter are compared with
the oracle in its BF16 mode (the same BF16 rounding points, FP64
accumulation): the remaining difference is FP32 vs FP64 accumulation and the
odd BF16 tie.  Training must fit a smooth target, and on a rendered scene the
trained network must beat the same quarter-resolution buffer upsampled
bilinearly (the method without the network).
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

from voxel_cone_tracing_amd import _lib as L
from voxel_cone_tracing_amd.context import Context
from voxel_cone_tracing_amd.scene import look_at, perspective

from helpers import SceneSpec

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _params(seed=0, scale=0.2):
    rng = np.random.default_rng(seed)
    p = (rng.standard_normal(L.MLP_PARAMS) * scale).astype(np.float32)
    p[34 * 64:34 * 64 + 64] = np.abs(p[34 * 64:34 * 64 + 64])
    return p


@pytest.mark.parametrize("n", [1, 100, 4096])
def test_forward_matches_oracle(gpu_available, orc, n):
    say(f"mlp forward n={n}")
    rng = np.random.default_rng(n)
    x = rng.random((n, 34)).astype(np.float32)
    p = _params(1)
    with Context(16, 16) as ctx:
        ctx.mlp_set_params(p)
        y = ctx.mlp_forward(x)
    want = orc.mlp_forward(p, x, bf16=True)
    err = np.abs(y - want)
    say(f"  max-abs {err.max():.2e}, max |y| {np.abs(want).max():.2f}")
    assert np.all(err <= 2e-2 + 2e-2 * np.abs(want))
    assert np.abs(y - want).mean() < 2e-3 * max(1.0, np.abs(want).mean())


@pytest.mark.parametrize("n", [64, 1000, 8192])
def test_gradient_matches_oracle(gpu_available, orc, n):
    say(f"mlp gradient n={n}")
    rng = np.random.default_rng(100 + n)
    x = rng.random((n, 34)).astype(np.float32)
    a = rng.random((n, 3)).astype(np.float32)
    t = rng.random((n, 3)).astype(np.float32)
    p = _params(2)
    with Context(16, 16) as ctx:
        ctx.mlp_set_params(p)
        loss, g = ctx.mlp_loss_grad(x, a, t)
        loss2, g2 = ctx.mlp_loss_grad(x, a, t)
    assert loss == loss2 and np.array_equal(g, g2), "gradient reduction must be deterministic"
    wl, wg = orc.mlp_loss_grad(p, x, a, t, bf16=True)
    rel = np.linalg.norm(g - wg) / np.linalg.norm(wg)
    say(f"  loss {loss:.6f} vs {wl:.6f}; gradient relative error {rel:.2e}")
    assert abs(loss - wl) <= 1e-3 * wl
    assert rel < 2e-2
    # every block of the parameter vector, not only the largest ones
    for lo, hi in [(0, 2176), (2176, 2240), (2240, 6336), (6336, 6400), (6400, 6592), (6592, 6595)]:
        r = np.linalg.norm(g[lo:hi] - wg[lo:hi]) / max(np.linalg.norm(wg[lo:hi]), 1e-12)
        assert r < 5e-2, (lo, hi, r)


def test_training_fits_a_smooth_target(gpu_available):
    rng = np.random.default_rng(7)
    n = 16384
    x = rng.random((n, 34)).astype(np.float32)
    t = np.stack([np.sin(3 * x[:, 0]) * x[:, 1], x[:, 2] * x[:, 3] + 0.5 * x[:, 4], np.cos(2 * x[:, 5])],
                 1).astype(np.float32)
    a = np.ones((n, 3), np.float32)
    with Context(16, 16) as ctx:
        ctx.mlp_init(3)
        l0, _ = ctx.mlp_loss_grad(x, a, t)
        ctx.upsample_set_samples(x, a, t)
        ctx.upsample_train(400, 4096, 3e-3, seed=1)  # device Adam
        l1, _ = ctx.mlp_loss_grad(x, a, t)
    say(f"loss {l0:.4f} -> {l1:.5f}")
    assert l1 < 0.1 * l0


def _soup_ctx(W, H, ntri=30000, res=64, mu=-4.0):
    sc = SceneSpec("soup", res=res, ntri=ntri, W=W, H=H, mu=mu)
    ctx = Context(W, H)
    sc.build(ctx, L.FP16)
    return sc, ctx


def _poses(k, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(k):
        d = rng.standard_normal(3)
        d /= np.linalg.norm(d)
        eye = d * rng.uniform(2.6, 3.6)
        out.append(look_at(eye, rng.uniform(-0.2, 0.2, 3), (0, 1, 0)))
    return out


def _plain(ctx, sc):
    sc_q(ctx, sc)
    ctx.render()
    return ctx.download_image()


def sc_q(ctx, sc):
    ctx.clear_queues()
    ctx.queue_model(sc.model)
    for (p, c, i, k) in sc.lights:
        ctx.queue_light(p, c, i, k)


def _bilinear(low, gn_alb, W, H):
    """The same quarter-resolution buffer upsampled without the network:
    bilinear in low-res coordinates, times albedo."""
    wl, hl = low.shape[1], low.shape[0]
    px = np.arange(W)
    py = np.arange(H)
    u = np.clip((px - 1) * 0.25, 0, wl - 1)
    v = np.clip((py - 1) * 0.25, 0, hl - 1)
    u0 = np.floor(u).astype(int)
    v0 = np.floor(v).astype(int)
    u1 = np.minimum(u0 + 1, wl - 1)
    v1 = np.minimum(v0 + 1, hl - 1)
    au = (u - u0)[None, :, None]
    av = (v - v0)[:, None, None]
    rgb = low[..., :3]
    top = rgb[v0][:, u0] * (1 - au) + rgb[v0][:, u1] * au
    bot = rgb[v1][:, u0] * (1 - au) + rgb[v1][:, u1] * au
    return (top * (1 - av) + bot * av) * gn_alb


def test_upsampler_trained_in_place_beats_bilinear_and_holds_40db(gpu_available):
    """Collect and train on the device end to end on a small scene, then the
    learned frame (per-tile fallback off) against bilinear upsampling of the
    same quarter-resolution buffer, and the per-tile fallback at 40 dB.

    The gain of the network over the surface-guided interpolation alone is
    tested at the benchmark workload with the committed network
    (test_c4_network_beats_guided_interpolation).  A network trained here in
    1,500 steps on 24 poses of this 4,000-triangle scene does not beat the
    interpolation on a held-out pose (measured 34.88 against 35.18 dB); the
    bound below only keeps it from falling further behind."""
    W, H = 256, 192
    # triangles of ~e^-2.5 = 0.08 units (several quarter-resolution samples
    # each at this viewport), so that most covered pixels are interpolated
    sc, ctx = _soup_ctx(W, H, ntri=4000, mu=-2.5)
    try:
        ctx.mlp_init(11)
        for i, view in enumerate(_poses(24, 5)):
            ctx.set_camera(view, sc.proj)
            sc_q(ctx, sc)
            ctx.voxelize()
            ctx.mip()
            ctx.gbuffer()
            ctx.upsample_collect(8192, 1000 + i)
        say(f"samples {ctx.upsample_count()}")
        loss = ctx.upsample_train(1500, 8192, 2e-3, seed=3)
        say(f"final training loss {loss:.3e}")
        trained = ctx.mlp_params()
        view = _poses(1, 99)[0]  # held-out pose
        ctx.set_camera(view, sc.proj)
        plain = _plain(ctx, sc)
        ctx.set_fallback(0.0)
        sc_q(ctx, sc)
        ctx.render_learned()
        learned = ctx.download_image()
        traced = ctx.traced_pixels()
        low = ctx.download_lowres()
        gp, gn = ctx.download_gbuffer()
        ctx.mlp_set_params(np.zeros_like(trained))
        sc_q(ctx, sc)
        ctx.render_learned()
        guided = ctx.download_image()
        ctx.mlp_set_params(trained)
        ctx.set_fallback(40.0)
        sc_q(ctx, sc)
        ctx.render_learned()
        checked = ctx.download_image()
        tiles, fb = ctx.fallback_stats()
    finally:
        ctx.close()
    cov = gp[..., 3] != 0
    mats = sc.mat_records
    mid = gn[..., 3].copy().view(np.uint32)
    alb = mats[np.where(cov, mid, 0)][..., 4:7]
    bil = _bilinear(low, alb, W, H)

    def psnr(img):
        e = img[..., :3][cov].astype(np.float64) - plain[..., :3][cov]
        return 10 * np.log10(1.0 / np.mean(e ** 2))

    p_l, p_g, p_b, p_c = psnr(learned), psnr(guided), psnr(bil), psnr(checked)
    say(f"PSNR vs plain full-res: learned {p_l:.2f} dB, guided interpolation {p_g:.2f} dB, bilinear {p_b:.2f} dB; "
        f"traced pixels {traced}/{int(cov.sum())}; with tile fallback {p_c:.2f} dB ({fb}/{tiles} tiles)")
    assert np.isfinite(learned).all()
    assert traced < 0.5 * cov.sum()  # the comparison is about interpolated pixels
    assert p_l > p_g - 0.5  # regression bound only (see the docstring)
    assert p_l > p_b + 1.0
    assert p_c >= 40.0


def _guided_numpy(gp, gn, low, alb, W, H):
    """Restatement of the surface-guided interpolation (the learned frame with a
    zero network): returns (albedo * interpolation, status) per pixel, status
    0 background, 1 interpolated, 2 traced at full resolution."""
    hl, wl = low.shape[:2]
    out = np.zeros((H, W, 3), np.float64)
    st = np.zeros((H, W), np.int8)
    lpos = gp[1::4, 1::4][:hl, :wl]
    lnrm = gn[1::4, 1::4][:hl, :wl]
    lmat = lnrm[..., 3].copy().view(np.uint32)
    mat = gn[..., 3].copy().view(np.uint32)
    for py in range(H):
        v = (py - 1) * 0.25
        jc = int(np.floor(v + 0.5))
        for px in range(W):
            if gp[py, px, 3] == 0:
                continue
            u = (px - 1) * 0.25
            ic = int(np.floor(u + 0.5))
            n = gn[py, px, :3].astype(np.float64)
            n = n / np.linalg.norm(n)
            wsum, acc = 0.0, np.zeros(3)
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    i, j = ic + dx, jc + dy
                    if not (0 <= i < wl and 0 <= j < hl):
                        continue
                    if lmat[j, i] != mat[py, px] or not low[j, i, 3] > 0:
                        continue
                    ln = lnrm[j, i, :3].astype(np.float64)
                    if not n.dot(ln) > 0.95 * np.linalg.norm(ln):
                        continue
                    if not abs(low[j, i, 3] - gp[py, px, 3]) < 0.05 * gp[py, px, 3]:
                        continue
                    wk = max(0.0, 1 - 0.5 * abs(u - i)) * max(0.0, 1 - 0.5 * abs(v - j))
                    wsum += wk
                    acc += wk * low[j, i, :3]
            if wsum > 0:
                st[py, px] = 1
                out[py, px] = alb[py, px] * acc / wsum
            else:
                st[py, px] = 2
    return out, st


def test_guided_interpolation_and_per_pixel_tracing(gpu_available):
    """With a zero network the learned frame is the surface-guided
    interpolation of the quarter-resolution trace (checked against a numpy
    restatement), pixels without a same-surface sample are traced at full
    resolution (bit-identical to the plain frame), and the listed count
    matches."""
    W, H = 96, 64
    sc, ctx = _soup_ctx(W, H, ntri=20000)
    try:
        ctx.mlp_set_params(np.zeros(L.MLP_PARAMS, np.float32))
        ctx.set_fallback(0.0)
        plain = _plain(ctx, sc)
        sc_q(ctx, sc)
        ctx.render_learned()
        img = ctx.download_image()
        traced = ctx.traced_pixels()
        low = ctx.download_lowres()
        gp, gn = ctx.download_gbuffer()
    finally:
        ctx.close()
    cov = gp[..., 3] != 0
    mid = gn[..., 3].copy().view(np.uint32)
    alb = sc.mat_records[np.where(cov, mid, 0)][..., 4:7]
    want, st = _guided_numpy(gp, gn, low, alb, W, H)
    say(f"interpolated {int((st == 1).sum())}, traced {int((st == 2).sum())} (GPU listed {traced})")
    assert traced == int((st == 2).sum())
    t = st == 2
    assert np.array_equal(img[t].view(np.uint32), plain[t].view(np.uint32))
    yy, xx = np.nonzero(st == 1)
    exact = ((yy & 3) == 1) & ((xx & 3) == 1)  # traced at quarter resolution: kept
    got = img[yy, xx, :3].astype(np.float64)
    err = np.abs(got[~exact] - want[yy[~exact], xx[~exact]])
    assert err.max() < 1e-4, err.max()
    assert np.allclose(got[exact], plain[yy[exact], xx[exact], :3], atol=1e-5)


@pytest.mark.parametrize("half", [True, False])
def test_tile_check_and_refinement_hold_40db(gpu_available, half):
    """The learned frame with the per-tile check at 40 dB: against the plain
    frame the result holds 40 dB, with (half=True) or without the
    half-resolution stage; the counters are consistent."""
    W, H = 320, 192
    sc, ctx = _soup_ctx(W, H, ntri=20000)
    try:
        ctx.mlp_set_params(np.zeros(L.MLP_PARAMS, np.float32))  # guided interpolation alone
        ctx.set_fallback(40.0)
        ctx.set_refine(half)
        plain = _plain(ctx, sc)
        sc_q(ctx, sc)
        ctx.render_learned()
        img = ctx.download_image()
        st = ctx.upsample_stats()
        gp, _ = ctx.download_gbuffer()
    finally:
        ctx.close()
    cov = gp[..., 3] != 0
    e = img[..., :3][cov].astype(np.float64) - plain[..., :3][cov]
    p = 10 * np.log10(1.0 / np.mean(e ** 2))
    say(f"half={half}: PSNR {p:.2f} dB, {st}")
    assert np.isfinite(img).all() and (img[..., 3] >= 0).all()  # no pixel left unresolved
    assert p >= 40.0
    assert st["tiles"] == 10 * 6 and 0 <= st["tiles_full"] <= st["tiles_refined"] <= st["tiles"]
    assert st["probes"] >= 16 * st["tiles"]
