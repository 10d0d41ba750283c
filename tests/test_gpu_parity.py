"""GPU parity of the three passes against the CPU oracle (through the C ABI).

Bit-exact: voxel occupancy, coordinates and stored level-0 values (both
averaging modes), every mip level of every direction map, G-buffer coverage,
positions, normals and material ids.  Tolerance: the cone-traced image
(FP32 on both sides; max-abs stated per test) and its PSNR against the FP64
ideal oracle.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

from voxel_cone_tracing_amd import _lib as L
from voxel_cone_tracing_amd.context import Context

from helpers import SceneSpec, bits, decode_level, encode_level

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


# ---- mip (kernel b) ---------------------------------------------------------------
def _random_level0(rng, res, dtype):
    n = res ** 3
    if dtype == L.RGBA8:
        v = rng.integers(0, 2 ** 32, size=n, dtype=np.uint64).astype(np.uint32)
        # a share of empty and of opaque voxels, as voxelisation produces
        v[rng.random(n) < 0.3] = 0
        op = rng.random(n) < 0.3
        v[op] = (v[op] & 0x00FFFFFF) | 0xFF000000
        return v.reshape(res, res, res)
    f = rng.random((n, 4)).astype(np.float32)
    f[rng.random(n) < 0.3] = 0
    f[rng.random(n) < 0.3, 3] = 1.0
    return encode_level(f.reshape(res, res, res, 4), dtype)


@pytest.mark.parametrize("dtype", [L.RGBA8, L.FP16, L.BF16, L.FP32])
@pytest.mark.parametrize("res,levels", [(32, 6), (64, 7), (256, 7)])
def test_mip_bit_exact(gpu_available, orc, dtype, res, levels):
    say(f"mip parity dtype={dtype} res={res} levels={levels}")
    rng = np.random.default_rng(res * 10 + dtype)
    l0 = _random_level0(rng, res, dtype)
    with Context(16, 16) as ctx:
        ctx.set_grid(res, levels, dtype)
        ctx.upload_level(0, 0, l0)
        ctx.mip()
        ctx.sync()
        ref = orc.mip_f32(res, levels, dtype, decode_level(l0, dtype))
        for l in range(1, levels):
            for d in range(6):
                got = ctx.download_level(l, d)
                want = encode_level(ref[(l, d)], dtype)
                assert np.array_equal(bits(got), bits(want)), f"level {l} dir {d} differs"


# ---- voxelize (kernel a) ----------------------------------------------------------
@pytest.mark.parametrize("kind,res,ntri", [("cornell", 64, 34), ("soup", 64, 3000), ("soup", 128, 20000)])
def test_voxelize_reference_average_bit_exact(gpu_available, orc, kind, res, ntri):
    say(f"voxelize reference average kind={kind} res={res}")
    sc = SceneSpec(kind, res=res, ntri=ntri)
    with Context(32, 32) as ctx:
        sc.build(ctx, L.RGBA8, L.AVG_REFERENCE)
        ctx.voxelize()
        got = ctx.download_level(0)
        nfrag, ncand = ctx.voxelize_stats()
    want = orc.voxelize(res, sc.cube_size, sc.M, sc.pos, sc.nrm, sc.tri_mat, sc.mat_records, sc.lights,
                        want=("ref", "count"))
    assert nfrag == want["fragments"]
    occ_g, occ_w = got != 0, want["ref"] != 0
    assert np.array_equal(occ_g, occ_w), f"occupancy differs at {np.argwhere(occ_g != occ_w)[:5]}"
    assert np.array_equal(got, want["ref"]), f"{np.count_nonzero(got != want['ref'])} voxels differ"


def test_voxelize_reference_average_many_models(gpu_available, orc):
    """37 queued models: the texels equal the oracle's for the same triangles
    as one list, since the canonical order runs over the queue."""
    say("voxelize reference average over 37 models")
    sc = SceneSpec("soup", res=64, ntri=3000)
    cuts = np.linspace(0, 3000, 38).astype(int)
    with Context(32, 32) as ctx:
        sc.build(ctx, L.RGBA8, L.AVG_REFERENCE)
        ctx.clear_queues()
        for a, b in zip(cuts[:-1], cuts[1:]):
            ctx.queue_model(ctx.model_create(sc.pos[a:b], sc.nrm[a:b], sc.tri_mat[a:b]))
        for (p, c, i, k) in sc.lights:
            ctx.queue_light(p, c, i, k)
        for _ in range(2):  # the second frame checks that the per-voxel counts start again from zero
            ctx.voxelize()
            got = ctx.download_level(0)
        nfrag, _ = ctx.voxelize_stats()
    want = orc.voxelize(sc.res, sc.cube_size, sc.M, sc.pos, sc.nrm, sc.tri_mat, sc.mat_records, sc.lights,
                        want=("ref", "count"))
    assert nfrag == want["fragments"]
    assert np.array_equal(got, want["ref"]), f"{np.count_nonzero(got != want['ref'])} voxels differ"


@pytest.mark.parametrize("dtype", [L.FP32, L.FP16, L.BF16, L.RGBA8])
@pytest.mark.parametrize("kind,res,ntri", [("cornell", 64, 34), ("soup", 128, 20000), ("cornell", 8, 34),
                                           ("soup", 256, 60000)])
def test_voxelize_exact_mean_bit_exact(gpu_available, orc, dtype, kind, res, ntri):
    say(f"voxelize exact mean dtype={dtype} kind={kind} res={res}")
    sc = SceneSpec(kind, res=res, ntri=ntri)
    with Context(32, 32) as ctx:
        sc.build(ctx, dtype, L.AVG_EXACT)
        for _ in range(2):  # the second frame checks that the sums were re-zeroed
            ctx.voxelize()
            got = ctx.download_level(0)
    want = orc.voxelize(res, sc.cube_size, sc.M, sc.pos, sc.nrm, sc.tri_mat, sc.mat_records, sc.lights,
                        want=("q32", "f64", "count"))
    w = encode_level(want["q32"], dtype)
    assert np.array_equal(bits(got), bits(w)), "level 0 differs"
    # error of the stored radiance against the FP64 exact mean
    err = np.abs(decode_level(got, dtype).astype(np.float64) - want["f64"]).max()
    bound = {L.FP32: 2e-5, L.FP16: 1e-3, L.BF16: 4.0e-3, L.RGBA8: 2.0e-3}[dtype]
    say(f"  max-abs vs FP64 = {err:.3e}")
    assert err <= bound


def test_voxelize_lists_past_their_capacity_bit_exact(gpu_available, orc):
    """Brick lists sized from an earlier frame's (triangle, brick) count: a
    frame whose geometry outgrew them (the same triangles, grown tenfold by
    the model matrix) has its overflowing bricks walk every triangle's brick
    box and stays bit-exact, and so does the frame after it (sized from the
    grown count)."""
    say("voxelize lists past their capacity")
    res = 256
    sc = SceneSpec("cornell", res=res)
    ntri = sc.pos.reshape(-1, 9).shape[0]
    small = np.diag([0.1, 0.1, 0.1, 1.0]).astype(np.float32)
    with Context(32, 32) as ctx:
        sc.build(ctx, L.FP16, L.AVG_EXACT)
        ctx.set_model_transform(sc.model, small)
        ctx.voxelize()  # the first frame waits for its count: the lists are sized from it
        _, small_pairs = ctx.voxelize_stats()
        ctx.set_model_transform(sc.model, sc.M)
        got = []
        for _ in range(2):
            ctx.voxelize()
            got.append(ctx.download_level(0))
        _, pairs = ctx.voxelize_stats()
    cap = max(small_pairs + small_pairs // 4, 2 * ntri + 1024)
    say(f"  pairs: shrunk box {small_pairs}, full box {pairs}, capacity of the second frame {cap}")
    assert pairs > cap
    want = orc.voxelize(res, sc.cube_size, sc.M, sc.pos, sc.nrm, sc.tri_mat, sc.mat_records, sc.lights,
                        want=("q32",))
    w = encode_level(want["q32"], L.FP16)
    for k, g in enumerate(got):
        assert np.array_equal(bits(g), bits(w)), f"frame {k + 2}: level 0 differs"


@pytest.mark.parametrize("dtype,average", [(L.FP16, L.AVG_EXACT), (L.RGBA8, L.AVG_REFERENCE)])
def test_voxelize_follows_material_changes_bit_exact(gpu_available, orc, dtype, average):
    """The set-up keeps the queued models' per-triangle material terms from
    frame to frame; a material upload (upload_material_data), a model's
    material assignment (set_model_material) and another model in the queue
    with the same triangle count each rebuild them, so every frame matches
    the oracle for the materials then in force."""
    import ctypes as C

    say(f"voxelize after material changes dtype={dtype} average={average}")
    sc = SceneSpec("cornell", res=64)
    recs = sc.mat_records.copy()
    with Context(32, 32) as ctx:
        sc.build(ctx, dtype, average)

        def check(tri_mat, tag):
            ctx.voxelize()
            got = ctx.download_level(0)
            want = orc.voxelize(sc.res, sc.cube_size, sc.M, sc.pos, sc.nrm, tri_mat, recs, sc.lights,
                                want=("ref", "q32"))
            w = want["ref"] if average == L.AVG_REFERENCE else encode_level(want["q32"], dtype)
            assert np.array_equal(bits(got), bits(w)), f"{tag}: level 0 differs"

        check(sc.tri_mat, "first frame")
        check(sc.tri_mat, "second frame")
        # a wall (material 2) turns green, emissive and transmissive
        r = recs[2].copy()
        r[4:7] = (0.1, 0.7, 0.3)
        r[12:15] = (0.5, 0.6, 0.7)
        r[16:19] = (0.2, 0.0, 0.05)
        r[21] = 0.4
        r[22] = np.array([4], np.int32).view(np.float32)[0]
        recs[2] = r
        m = L.Material.from_buffer_copy(np.ascontiguousarray(r).tobytes())
        L.check(ctx.lib.vct_upload_material_data(ctx.h, C.byref(m), 2))
        check(sc.tri_mat, "after the upload")
        check(sc.tri_mat, "the frame after the upload")
        ctx.set_model_material(2, sc.model)
        every2 = np.full_like(sc.tri_mat, 2)
        check(every2, "after set_model_material")
        # another model with the same triangles and count, in the original materials
        other = ctx.model_create(sc.pos, sc.nrm, sc.tri_mat)
        ctx.set_model_transform(other, sc.M)
        ctx.clear_queues()
        ctx.queue_model(other)
        for (p, c, i, k) in sc.lights:
            ctx.queue_light(p, c, i, k)
        check(sc.tri_mat, "another model queued")
        ctx.clear_queues()
        ctx.queue_model(sc.model)
        for (p, c, i, k) in sc.lights:
            ctx.queue_light(p, c, i, k)
        check(every2, "the first model queued again")


# ---- voxelize under a model matrix (voxelize.vert:24-30) ----------------------------
@pytest.mark.parametrize("dtype,average", [(L.RGBA8, L.AVG_REFERENCE), (L.FP16, L.AVG_EXACT), (L.FP32, L.AVG_EXACT)])
def test_voxelize_with_model_matrix_bit_exact(gpu_available, orc, dtype, average):
    """Rotation, translation and non-uniform scale: world positions through
    the model matrix, normals through its inverse transpose."""
    say(f"voxelize with a model matrix dtype={dtype} average={average}")
    sc = SceneSpec("soup", res=128, ntri=20000, model="rot_scale")
    with Context(32, 32) as ctx:
        sc.build(ctx, dtype, average)
        ctx.voxelize()
        got = ctx.download_level(0)
    want = orc.voxelize(sc.res, sc.cube_size, sc.M, sc.pos, sc.nrm, sc.tri_mat, sc.mat_records, sc.lights,
                        want=("ref", "q32"))
    w = want["ref"] if average == L.AVG_REFERENCE else encode_level(want["q32"], dtype)
    assert np.array_equal(bits(got), bits(w)), "level 0 differs"


@pytest.mark.parametrize("dtype,average", [(L.RGBA8, L.AVG_REFERENCE), (L.FP16, L.AVG_EXACT)])
def test_voxelize_many_models_bit_exact(gpu_available, orc, dtype, average):
    """Any number of queued models (renderer.cpp:242-256): 37 draws, two of
    them with a model matrix, in one frame; two frames in a row."""
    say(f"voxelize over 37 models dtype={dtype} average={average}")
    from helpers import model_matrix

    sc = SceneSpec("soup", res=128, ntri=20000)
    cuts = np.linspace(0, 20000, 38).astype(int)
    draws = []
    with Context(32, 32) as ctx:
        sc.build(ctx, dtype, average)
        ctx.clear_queues()
        for k, (a, b) in enumerate(zip(cuts[:-1], cuts[1:])):
            m = ctx.model_create(sc.pos[a:b], sc.nrm[a:b], sc.tri_mat[a:b])
            M = model_matrix("rot_scale") if k in (3, 30) else sc.M
            ctx.set_model_transform(m, M)
            ctx.queue_model(m)
            draws.append((M, sc.pos[a:b], sc.nrm[a:b], sc.tri_mat[a:b]))
        for (p, c, i, k) in sc.lights:
            ctx.queue_light(p, c, i, k)
        for _ in range(2):  # the second frame reuses the draw table
            ctx.voxelize()
        got = ctx.download_level(0)
    want = orc.voxelize_draws(sc.res, sc.cube_size, draws, sc.mat_records, sc.lights, want=("ref", "q32"))
    w = want["ref"] if average == L.AVG_REFERENCE else encode_level(want["q32"], dtype)
    assert np.array_equal(bits(got), bits(w)), "level 0 differs"


# ---- G-buffer -----------------------------------------------------------------------
INSIDE = dict(eye=(0.12, 0.05, 0.2), target=(-0.25, -0.1, -0.5), fovy_deg=75.0)


@pytest.mark.parametrize("kind,ntri,W,H,extra", [
    ("cornell", 34, 96, 64, {}), ("soup", 5000, 160, 90, {}), ("cornell", 34, 512, 512, {}),
    ("soup", 40000, 640, 360, {}), ("soup", 3000, 100, 70, {}),
    # the camera inside the box: walls crossing the near plane are clipped and drawn as fans
    ("cornell", 34, 128, 96, INSIDE), ("cornell", 34, 640, 480, INSIDE),
    # a camera among the soup's triangles: many cross the near plane
    ("soup", 20000, 320, 180, dict(eye=(0.05, 0.02, 0.1), target=(0.4, -0.1, -1.0), fovy_deg=90.0)),
    # a model matrix (voxel_cone_tracing.vert:22-28)
    ("soup", 20000, 160, 90, dict(model="rot_scale"))],
    ids=["cornell", "soup", "cornell512", "soup640", "ragged", "inside", "inside640", "in_soup", "model"])
def test_gbuffer_bit_exact(gpu_available, orc, kind, ntri, W, H, extra):
    say(f"gbuffer parity kind={kind} {W}x{H} {extra}")
    sc = SceneSpec(kind, ntri=ntri, W=W, H=H, **extra)
    with Context(W, H) as ctx:
        sc.build(ctx, L.FP16)
        for _ in range(2):  # the second frame reuses the tile lists and counters
            ctx.gbuffer()
        gp, gn = ctx.download_gbuffer()
    p32, n32, _, _, keys = orc.gbuffer(W, H, sc.view, sc.proj, [(sc.M, sc.pos, sc.nrm, sc.tri_mat)])
    cov_g, cov_w = gp[..., 3] != 0, p32[..., 3] != 0
    assert cov_w.any()
    fans = int(((keys & np.uint64(7)) != 0).sum())
    say(f"  covered {int(cov_w.sum())} pixels, {fans} of them by fan triangles beyond the first")
    assert np.array_equal(cov_g, cov_w)
    assert np.array_equal(bits(gp), bits(p32))
    assert np.array_equal(bits(gn), bits(n32))


def test_gbuffer_many_models(gpu_available, orc):
    """Any number of queued models (renderer.cpp:242-256): 37 draws, two of
    them with a model matrix, camera inside the soup (clipped triangles)."""
    say("gbuffer parity over 37 models")
    from helpers import model_matrix

    sc = SceneSpec("soup", ntri=6000, W=200, H=120, eye=(0.05, 0.02, 0.6), target=(0.0, 0.0, -1.0), fovy_deg=80.0)
    cuts = np.linspace(0, 6000, 38).astype(int)
    draws = []
    with Context(sc.W, sc.H) as ctx:
        sc.build(ctx, L.FP16)
        ctx.clear_queues()
        for k, (a, b) in enumerate(zip(cuts[:-1], cuts[1:])):
            m = ctx.model_create(sc.pos[a:b], sc.nrm[a:b], sc.tri_mat[a:b])
            M = model_matrix("rot_scale") if k in (3, 30) else sc.M
            ctx.set_model_transform(m, M)
            ctx.queue_model(m)
            draws.append((M, sc.pos[a:b], sc.nrm[a:b], sc.tri_mat[a:b]))
        for _ in range(2):  # the second frame reuses the draw table and the lists
            ctx.gbuffer()
        gp, gn = ctx.download_gbuffer()
    p32, n32, _, _, _ = orc.gbuffer(sc.W, sc.H, sc.view, sc.proj, draws)
    assert (p32[..., 3] != 0).any()
    assert np.array_equal(bits(gp), bits(p32))
    assert np.array_equal(bits(gn), bits(n32))


# ---- trace (kernel c) ----------------------------------------------------------------
@pytest.mark.parametrize("kind,dtype,res,ntri", [("cornell", L.RGBA8, 64, 34), ("cornell", L.FP16, 64, 34),
                                                 ("soup", L.FP16, 64, 4000), ("soup", L.BF16, 64, 4000),
                                                 ("soup", L.FP32, 64, 4000)])
def test_trace_matches_oracle(gpu_available, orc, kind, dtype, res, ntri):
    say(f"trace parity kind={kind} dtype={dtype}")
    W, H = 96, 64
    sc = SceneSpec(kind, res=res, ntri=ntri, W=W, H=H)
    levels = 7
    with Context(W, H) as ctx:
        sc.build(ctx, dtype)
        ctx.voxelize()
        ctx.mip()
        ctx.gbuffer()
        ctx.trace()
        img = ctx.download_image()
        gp, gn = ctx.download_gbuffer()
        l0 = decode_level(ctx.download_level(0), dtype)
        mips = {(l, d): decode_level(ctx.download_level(l, d), dtype) for l in range(1, levels) for d in range(6)}
    rng = np.random.default_rng(7)
    pix = np.sort(rng.choice(W * H, size=min(W * H, 1500), replace=False)).astype(np.int64)
    want, _ = orc.trace(np.float32, res, levels, l0, mips, gp, gn, sc.mat_records, sc.lights, sc.cam, sc.cube_size,
                        cone_set=sc.cone_set, pixels=pix)
    got = img.reshape(-1, 4)[pix]
    err = np.abs(got[:, :3] - want[:, :3]).max()
    say(f"  max-abs vs FP32 oracle = {err:.3e}")
    assert err <= 2e-3


@pytest.mark.parametrize("case", ["cones16", "lights3"])
def test_trace_full_block_shape_matches_oracle(gpu_available, orc, case):
    """Configurations beyond the 128-pixel block shape (more than 9 diffuse
    cones, or shadow cones of more than 2 lights) run the 64-pixel shape with
    every LDS slot; same oracle bar as the other trace tests."""
    say(f"trace parity, full block shape: {case}")
    W, H, res = 96, 64, 64
    if case == "cones16":
        sc = SceneSpec("soup", res=res, ntri=4000, W=W, H=H)
        sc.cone_set = L.CONES_16
    else:
        sc = SceneSpec("cornell", res=res, W=W, H=H, glass=True, phases=(True, True, True, True))
        sc.lights = sc.lights + [((0.3, 0.2, 0.35), (0.9, 0.6, 0.4), 0.7, L.LIGHT_POINT),
                                 ((-0.6, 1.0, 0.8), (0.5, 0.6, 1.0), 0.8, L.LIGHT_DIRECTIONAL)]
    with Context(W, H) as ctx:
        sc.build(ctx, L.FP16)
        ctx.render()
        img = ctx.download_image()
        gp, gn = ctx.download_gbuffer()
        l0 = decode_level(ctx.download_level(0), L.FP16)
        mips = {(l, d): decode_level(ctx.download_level(l, d), L.FP16) for l in range(1, 7) for d in range(6)}
    cov = gp[..., 3].reshape(-1) != 0
    pix = np.flatnonzero(cov).astype(np.int64)
    assert pix.size > 0
    want, _ = orc.trace(np.float32, res, 7, l0, mips, gp, gn, sc.mat_records, sc.lights, sc.cam, sc.cube_size,
                        cone_set=sc.cone_set, phases=sc.phases, pixels=pix)
    err = np.abs(img.reshape(-1, 4)[pix, :3] - want[:, :3]).max()
    say(f"  max-abs vs FP32 oracle = {err:.3e}")
    assert err <= 2e-3


@pytest.mark.parametrize("dtype", [L.RGBA8, L.FP16])
def test_full_compose_matches_oracle(gpu_available, orc, dtype):
    """All phases on (direct light + shadow cones, diffuse, specular and
    refraction cones, emission, the transmissive override), on the GPU's own
    volume and G-buffer."""
    say(f"full compose parity dtype={dtype}")
    W, H, res = 128, 96, 64
    sc = SceneSpec("cornell", res=res, W=W, H=H, glass=True, phases=(True, True, True, True))
    with Context(W, H) as ctx:
        sc.build(ctx, dtype)
        ctx.render()
        img = ctx.download_image()
        gp, gn = ctx.download_gbuffer()
        l0 = decode_level(ctx.download_level(0), dtype)
        mips = {(l, d): decode_level(ctx.download_level(l, d), dtype) for l in range(1, 7) for d in range(6)}
    cov = gp[..., 3].reshape(-1) != 0
    pix = np.flatnonzero(cov).astype(np.int64)
    want, _ = orc.trace(np.float32, res, 7, l0, mips, gp, gn, sc.mat_records, sc.lights, sc.cam, sc.cube_size,
                        cone_set=sc.cone_set, phases=sc.phases, pixels=pix)
    err = np.abs(img.reshape(-1, 4)[pix, :3] - want[:, :3]).max()
    say(f"  max-abs vs FP32 oracle = {err:.3e}")
    assert err <= 2e-3


def test_pass_timing_mask(gpu_available):
    """vct_set_pass_timing: with timing off the frame is the same bit for bit
    and every pass reads 0 ms; a pass left out of the mask reads 0 while the
    others are timed; a mask past pass 5 is an error."""
    say("pass timing mask")
    W, H, res = 96, 64, 64
    sc = SceneSpec("cornell", res=res, W=W, H=H)

    def frame(ctx):
        ctx.clear_queues()
        ctx.queue_model(sc.model)
        for (p, c, i, k) in sc.lights:
            ctx.queue_light(p, c, i, k)
        ctx.render()
        return ctx.download_image(), ctx.pass_times()

    with Context(W, H) as ctx:
        sc.build(ctx, L.FP16)
        img_on, t_on = frame(ctx)
        ctx.set_pass_timing(0)
        img_off, t_off = frame(ctx)
        ctx.set_pass_timing(0x3F & ~(1 << 3))  # all but the trace
        _, t_some = frame(ctx)
        with pytest.raises(L.VctError):
            ctx.set_pass_timing(0x40)
    say(f"  on {np.round(t_on[:4], 4)}, off {t_off[:4]}, no trace {np.round(t_some[:4], 4)}")
    assert np.array_equal(img_on, img_off)
    assert (t_on[:4] > 0).all() and (t_off == 0).all()
    assert t_some[3] == 0 and (t_some[[0, 1, 2]] > 0).all()


def test_trace_ragged_viewport_matches_oracle(gpu_available, orc):
    """A viewport whose block grid is ragged: 88x100 is 6 x 13 blocks of
    16x8 pixels, partial on the right and bottom edges, 8 block rows in the
    XCD-interleaved order and 5 left in dispatch order.  The Cornell scene
    covers every pixel, so a block traced twice or not at all shows up."""
    say("trace parity, ragged 88x100 viewport")
    W, H, res = 88, 100, 64
    sc = SceneSpec("cornell", res=res, W=W, H=H, glass=True, phases=(True, True, True, True))
    with Context(W, H) as ctx:
        sc.build(ctx, L.FP16)
        ctx.render()
        img = ctx.download_image()
        gp, gn = ctx.download_gbuffer()
        l0 = decode_level(ctx.download_level(0), L.FP16)
        mips = {(l, d): decode_level(ctx.download_level(l, d), L.FP16) for l in range(1, 7) for d in range(6)}
    cov = gp[..., 3].reshape(-1) != 0
    assert cov.all()
    pix = np.arange(W * H, dtype=np.int64)
    want, _ = orc.trace(np.float32, res, 7, l0, mips, gp, gn, sc.mat_records, sc.lights, sc.cam, sc.cube_size,
                        cone_set=sc.cone_set, phases=sc.phases, pixels=pix)
    err = np.abs(img.reshape(-1, 4)[:, :3] - want[:, :3]).max()
    say(f"  max-abs vs FP32 oracle = {err:.3e}")
    assert err <= 2e-3


@pytest.mark.parametrize("view_dir,lod,extra,dtype", [
    (0, 0.0, {}, L.RGBA8), (3, 1.5, {}, L.RGBA8), (5, 4.0, {}, L.FP16),
    # the camera inside the box: clipped walls drawn as fans, in draw order
    (1, 0.5, dict(eye=(0.12, 0.05, 0.2), target=(-0.25, -0.1, -0.5), fovy_deg=75.0), L.RGBA8)])
def test_voxel_view_matches_oracle(gpu_available, orc, view_dir, lod, extra, dtype):
    """The voxel-view debug path (voxel_cone_tracing.frag:253-257) with GL's
    in-order SRC_ALPHA blending of every fragment that passes the depth test
    (renderer.cpp:384-388), against the oracle's sequential draw."""
    say(f"voxel view dir={view_dir} lod={lod} {extra} dtype={dtype}")
    W, H, res = 96, 64, 64
    sc = SceneSpec("cornell", res=res, W=W, H=H, **extra)
    with Context(W, H) as ctx:
        sc.build(ctx, dtype)
        ctx.set_voxel_view(view_dir, lod)
        ctx.render()
        # a second frame reuses the sort buffers
        ctx.queue_model(sc.model)
        for (p, c, i, k) in sc.lights:
            ctx.queue_light(p, c, i, k)
        ctx.render()
        img = ctx.download_image()
        l0 = decode_level(ctx.download_level(0), dtype)
        mips = {(l, d): decode_level(ctx.download_level(l, d), dtype) for l in range(1, 7) for d in range(6)}
    want = orc.voxel_view(W, H, sc.view, sc.proj, [(sc.M, sc.pos)], res, 7, l0, mips, sc.cube_size, view_dir, lod)
    err = np.abs(img - want).max()
    say(f"  max-abs vs FP32 oracle = {err:.3e}")
    assert err == 0.0  # the same float operations in the same order (measured bit-exact)


def test_voxel_view_over_many_models(gpu_available, orc):
    """Draw order runs over the queued models (renderer.cpp:242-256): the
    Cornell box split into 5 models queued in reverse, two of them moved."""
    say("voxel view over 5 models")
    from helpers import model_matrix

    W, H, res = 96, 64, 64
    sc = SceneSpec("cornell", res=res, W=W, H=H)
    cuts = np.linspace(0, sc.pos.shape[0], 6).astype(int)
    draws = []
    with Context(W, H) as ctx:
        sc.build(ctx, L.RGBA8)
        ctx.clear_queues()
        models = []
        for k, (a, b) in enumerate(zip(cuts[:-1], cuts[1:])):
            m = ctx.model_create(sc.pos[a:b], sc.nrm[a:b], sc.tri_mat[a:b])
            M = model_matrix("rot_scale") if k in (1, 3) else sc.M
            ctx.set_model_transform(m, M)
            models.append((m, M, sc.pos[a:b]))
        for m, M, pos in reversed(models):
            ctx.queue_model(m)
            draws.append((M, pos))
        for (p, c, i, k) in sc.lights:
            ctx.queue_light(p, c, i, k)
        ctx.set_voxel_view(2, 0.7)
        ctx.render()
        img = ctx.download_image()
        l0 = decode_level(ctx.download_level(0), L.RGBA8)
        mips = {(l, d): decode_level(ctx.download_level(l, d), L.RGBA8) for l in range(1, 7) for d in range(6)}
    want = orc.voxel_view(W, H, sc.view, sc.proj, draws, res, 7, l0, mips, sc.cube_size, 2, 0.7)
    err = np.abs(img - want).max()
    say(f"  max-abs vs FP32 oracle = {err:.3e}")
    assert err == 0.0  # the same float operations in the same order (measured bit-exact)


@pytest.mark.parametrize("dtype", [L.RGBA8, L.FP16, L.BF16, L.FP32])
@pytest.mark.parametrize("view_dir,lod", [(1, 0.0), (2, 0.6), (4, 2.3)])
def test_sampler_at_the_volume_faces(gpu_available, orc, dtype, view_dir, lod):
    """The marcher's sampler (x-paired texels for 8- and 16-bit storage) on a
    random volume at points spread over the whole cube and up to half a texel
    past its faces, where the zero border and the first / last columns meet
    (the voxel view samples exactly one textureLod per pixel)."""
    say(f"sampler at faces dtype={dtype} dir={view_dir} lod={lod}")
    W, H, res, levels = 64, 32, 8, 4
    sc = SceneSpec("cornell", res=res, W=W, H=H)
    rng = np.random.default_rng(100 + dtype)
    with Context(W, H) as ctx:
        sc.build(ctx, dtype, levels=levels)
        want_levels = {}
        for lvl in range(levels):
            s = res >> lvl
            for d in range(1 if lvl == 0 else 6):
                f = rng.random((s, s, s, 4)).astype(np.float32)
                enc = encode_level(f, dtype)
                ctx.upload_level(lvl, d, enc)
                want_levels[(lvl, d)] = decode_level(enc, dtype)
        # biased coordinates in [-0.06, 0.999]: columns -1 and s-1 of every level are hit
        b = rng.uniform(-0.06, 0.999, (H, W, 3)).astype(np.float32)
        b[0, :, 0] = -0.5 / res + 1e-3  # the first texel's left border, level 0
        b[1, :, 0] = 1.0 - 0.5 / res + 1e-3  # past the last texel centre, level 0
        world = (b - 0.5) * 2.0 * sc.cube_size
        gp = np.concatenate([world, np.ones((H, W, 1), np.float32)], -1).astype(np.float32)
        gn = np.zeros((H, W, 4), np.float32)
        gn[..., 1] = 1.0
        gn[..., 3] = np.uint32(1).view(np.float32)
        ctx.upload_gbuffer(gp, gn)
        ctx.set_voxel_view(view_dir, lod)
        ctx.trace()
        img = ctx.download_image()
    l0 = want_levels[(0, 0)]
    mips = {(lv, d): want_levels[(lv, d)] for lv in range(1, levels) for d in range(6)}
    want, _ = orc.trace(np.float32, res, levels, l0, mips, gp, gn, sc.mat_records, sc.lights, sc.cam, sc.cube_size,
                        cone_set=sc.cone_set, phases=sc.phases, view_dir=view_dir, view_lod=lod)
    err = np.abs(img.reshape(-1, 4) - want).max()
    say(f"  max-abs vs FP32 oracle = {err:.3e}")
    # 16-bit pairs blend with weights rounded to the storage type (2^-11 / 2^-8)
    assert err <= {L.RGBA8: 1e-5, L.FP16: 2e-3, L.BF16: 1.6e-2, L.FP32: 1e-5}[dtype]
