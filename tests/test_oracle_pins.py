"""Pins of the CPU oracle (CPU only).

The reference ships no tests or golden vectors, so each oracle function is
pinned by what the mathematics fixes: known answers worked out by hand from
the shader text, closed forms on crafted volumes, invariants (symmetry,
watertight rasterisation), and library routines for the number formats.
Each check is chosen so that a plausible slip (a swapped pair, a dropped
term, a wrong tie rule, truncation instead of rounding) fails it.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle

# --------------------------------------------------------------------------------------
# imageAtomicRGBA8Avg (voxelize.frag:66-120)
# --------------------------------------------------------------------------------------


def unpack(v):
    return [(v >> (8 * k)) & 0xFF for k in range(4)]


def count_of(v):
    b = unpack(v)
    return (b[0] & 1) | ((b[1] & 1) << 1) | ((b[2] & 1) << 2) | ((b[3] & 1) << 3)


def test_first_fragment_truncates_and_sets_count_one():
    # val = 255 * (0.5, 0.25, 1, 1) = (127.5, 63.75, 255, 255): uint() truncates,
    # & 0xFE clears the count bits, the count 1 goes to bit 0 of R.
    s = oracle.rgba8_avg_step(0, [0.5, 0.25, 1.0, 1.0])
    assert unpack(s) == [127, 62, 254, 254]
    assert count_of(s) == 1
    assert s == 0xFEFE3E7F


def test_second_fragment_running_average_hand_computed():
    oracle.set_round_mode(False)
    s = oracle.rgba8_avg_step(0, [0.5, 0.25, 1.0, 1.0])  # stored colour (126, 62, 254, 254), n = 1
    s = oracle.rgba8_avg_step(s, [0.0, 0.0, 0.0, 1.0])   # val = (0, 0, 0, 255)
    # r = (126*1 + 0)/2 = 63 -> round(31.5)*2 = 64; (62)/2 = 31 -> round(15.5)*2 = 32;
    # 254/2 = 127 -> round(63.5)*2 = 128; (254+255)/2 = 254.5 -> round(127.25)*2 = 254; n = 2
    assert [b & 0xFE for b in unpack(s)] == [64, 32, 128, 254]
    assert count_of(s) == 2
    assert s == 0xFE802140


def test_round_tie_direction_is_a_switch():
    # r/2 = 32.5 exactly: ties-to-even gives 32 -> 64, ties-away gives 33 -> 66.
    first = [130 / 255.0, 0, 0, 1]  # stored R = 130 & 0xFE = 130
    second = [0.0, 0, 0, 1]          # (130 + 0) / 2 = 65 -> 65/2 = 32.5
    oracle.set_round_mode(False)
    s_even = oracle.rgba8_avg_step(oracle.rgba8_avg_step(0, first), second)
    oracle.set_round_mode(True)
    s_away = oracle.rgba8_avg_step(oracle.rgba8_avg_step(0, first), second)
    oracle.set_round_mode(False)
    assert unpack(s_even)[0] & 0xFE == 64
    assert unpack(s_away)[0] & 0xFE == 66


def test_count_nibble_wraps_after_sixteen_and_resets_the_mean():
    c = [0.2, 0.4, 0.6, 1.0]
    s = 0
    for k in range(1, 17):
        s = oracle.rgba8_avg_step(s, c)
        assert count_of(s) == k % 16
    # count 0 (wrapped): the next fragment replaces the colour: r = r*0 + val,
    # then round(val / 2) * 2; with alpha 255: round(127.5) * 2 = 256 -> & 0xFF = 0
    s = oracle.rgba8_avg_step(s, [1.0, 0.0, 0.0, 1.0])
    b = unpack(s)
    assert count_of(s) == 1
    assert b[0] & 0xFE == 0   # 255 -> 256 -> wraps to 0 in the byte
    assert b[3] == 0          # the alpha byte wraps as well (reference quirk)


def test_stored_zero_is_treated_as_empty():
    # a stored value of exactly 0 makes the first CAS succeed (voxelize.frag:106)
    assert oracle.rgba8_avg_step(0, [0.0, 0.0, 0.0, 0.0]) == 1  # colour 0, count 1 in bit 0


def test_identical_fragments_are_a_fixed_point():
    c = [100 / 255.0, 50 / 255.0, 200 / 255.0, 1.0]
    s = oracle.rgba8_avg_step(0, c)
    for k in range(2, 12):
        s = oracle.rgba8_avg_step(s, c)
        # the mean of identical even values stays at that value
        assert [b & 0xFE for b in unpack(s)] == [100, 50, 200, 254]


# --------------------------------------------------------------------------------------
# number formats: checked against numpy / torch conversion routines
# --------------------------------------------------------------------------------------


def test_half_encoding_matches_numpy_round_to_nearest_even():
    rng = np.random.default_rng(0)
    x = np.concatenate([
        rng.standard_normal(200000).astype(np.float32) * np.float32(10.0) ** rng.integers(-8, 5, 200000),
        np.array([0.0, -0.0, 65504.0, 65519.99, 65520.0, 6e-8, 3e-8, 5.96e-8, 1e-5, np.inf, -np.inf], np.float32),
    ]).astype(np.float32)
    got = oracle.encode_f16(x)
    want = x.astype(np.float16).view(np.uint16)
    assert np.array_equal(got, want)
    back = oracle.decode_f16(got)
    assert np.array_equal(back, want.view(np.float16).astype(np.float32))


def test_bf16_encoding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(1)
    x = (rng.standard_normal(200000) * 10.0 ** rng.integers(-30, 30, 200000)).astype(np.float32)
    got = oracle.encode_bf16(x)
    want = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(got, want)


def test_unorm8_is_c_over_255_and_round_trips():
    c = np.arange(256, dtype=np.uint32)
    packed = c | (c << 8) | (c << 16) | (c << 24)
    f = oracle.decode_rgba8(packed)
    assert np.array_equal(f[:, 0], (c.astype(np.float32) / np.float32(255.0)))
    assert np.array_equal(oracle.encode_rgba8(f), packed)
    # rounding to nearest, not truncation: 0.5/255 above a code rounds up
    v = np.array([[(10.6) / 255.0, (10.4) / 255.0, 2.0, -1.0]], np.float32)
    assert unpack(int(oracle.encode_rgba8(v)[0])) == [11, 10, 255, 0]


# --------------------------------------------------------------------------------------
# anisotropic mip (mipmap.comp)
# --------------------------------------------------------------------------------------


def test_mip_uniform_volume_closed_form():
    # alpha_blend(v, v) = v + (1 - a) v = (2 - a) v for every pair: level 1 = (2 - a) * v
    res = 8
    v = np.array([0.3, 0.2, 0.1, 0.25])
    l0 = np.broadcast_to(v, (res, res, res, 4)).copy()
    m = oracle.mip_f64(res, 3, l0)
    for d in range(6):
        assert np.allclose(m[(1, d)], (2 - 0.25) * v)
        w = (2 - 0.25) * v
        assert np.allclose(m[(2, d)], (2 - w[3]) * w)


def test_mip_front_hides_back_per_direction():
    # two opaque texels along each axis: the one nearer the viewer looking
    # along the map direction wins; the others of the 2x2x2 block are empty.
    A = np.array([1.0, 0.0, 0.0, 1.0])
    B = np.array([0.0, 1.0, 0.0, 1.0])
    for axis, (neg, pos) in enumerate([(0, 1), (2, 3), (4, 5)]):
        l0 = np.zeros((2, 2, 2, 4))
        lo = [0, 0, 0]
        hi = [0, 0, 0]
        hi[axis] = 1
        l0[lo[2], lo[1], lo[0]] = A  # coordinate 0 on the axis
        l0[hi[2], hi[1], hi[0]] = B  # coordinate 1 on the axis
        m = oracle.mip_f64(2, 2, l0)
        # map "-axis" looks toward -axis: front is the +1 texel (B); map "+axis": front is A
        assert np.allclose(m[(1, neg)][0, 0, 0], B / 4)
        assert np.allclose(m[(1, pos)][0, 0, 0], A / 4)
        for d in range(6):
            if d not in (neg, pos):
                assert np.allclose(m[(1, d)][0, 0, 0], (A + B) / 4)


def test_mip_mirror_symmetry_swaps_maps():
    rng = np.random.default_rng(3)
    res = 16
    l0 = rng.random((res, res, res, 4))
    l0[..., 3] = rng.random((res, res, res)) > 0.5
    m = oracle.mip_f64(res, 4, l0)
    mx = oracle.mip_f64(res, 4, l0[:, :, ::-1].copy())   # flip x
    my = oracle.mip_f64(res, 4, l0[:, ::-1, :].copy())   # flip y
    mz = oracle.mip_f64(res, 4, l0[::-1, :, :].copy())   # flip z
    for l in range(1, 4):
        assert np.allclose(mx[(l, 0)][:, :, ::-1], m[(l, 1)])
        assert np.allclose(mx[(l, 1)][:, :, ::-1], m[(l, 0)])
        assert np.allclose(mx[(l, 2)][:, :, ::-1], m[(l, 2)])
        assert np.allclose(my[(l, 2)][:, ::-1, :], m[(l, 3)])
        assert np.allclose(mz[(l, 4)][::-1, :, :], m[(l, 5)])


def test_mip_f32_is_f64_rounded_within_one_storage_step():
    rng = np.random.default_rng(4)
    res = 16
    l0 = oracle.store_round(rng.random((res, res, res, 4)).astype(np.float32), 0)
    m32 = oracle.mip_f32(res, 5, 0, l0)
    m64 = oracle.mip_f64(res, 5, l0.astype(np.float64))
    # level 1 is one unorm8 rounding (after the clamp to [0, 1]) away from the exact value
    assert np.abs(m32[(1, 0)] - np.clip(m64[(1, 0)], 0, 1)).max() <= 0.5 / 255 + 1e-6


# --------------------------------------------------------------------------------------
# textureLod: LINEAR_MIPMAP_LINEAR + CLAMP_TO_BORDER (texture_3d.cpp:10-15)
# --------------------------------------------------------------------------------------


def _vol(res=8, levels=4, seed=5):
    rng = np.random.default_rng(seed)
    l0 = rng.random((res, res, res, 4))
    lv = [l0]
    for l in range(1, levels):
        s = res >> l
        lv.append([rng.random((s, s, s, 4)) for _ in range(6)])
    return lv


def test_sampler_texel_centres_are_exact():
    lv = _vol()
    for (x, y, z) in [(0, 0, 0), (3, 5, 7), (7, 7, 7)]:
        p = [(x + 0.5) / 8, (y + 0.5) / 8, (z + 0.5) / 8]
        assert np.allclose(oracle.sample_lod(lv, 2, p, 0.0), lv[0][z, y, x])
    # level 1 texel centre at integer lod 1
    p = [(1 + 0.5) / 4, (2 + 0.5) / 4, (3 + 0.5) / 4]
    assert np.allclose(oracle.sample_lod(lv, 3, p, 1.0), lv[1][3][3, 2, 1])


def test_sampler_is_linear_between_centres_and_levels():
    lv = _vol()
    a, b = lv[0][2, 2, 2], lv[0][2, 2, 3]
    p = [3.0 / 8, 2.5 / 8, 2.5 / 8]  # half way between x = 2 and x = 3 centres
    assert np.allclose(oracle.sample_lod(lv, 0, p, 0.0), 0.5 * (a + b))
    p = [(1 + 0.5) / 4, (1 + 0.5) / 4, (1 + 0.5) / 4]
    s1 = oracle.sample_lod(lv, 4, p, 1.0)
    s2 = oracle.sample_lod(lv, 4, p, 2.0)
    assert np.allclose(oracle.sample_lod(lv, 4, p, 1.25), 0.75 * s1 + 0.25 * s2)


def test_sampler_border_is_zero_and_lod_clamps():
    lv = _vol()
    # a quarter texel outside x = 0: weight 0.25 on the zero border
    p = [0.25 / 8, 0.5 / 8, 0.5 / 8]
    assert np.allclose(oracle.sample_lod(lv, 1, p, 0.0), 0.75 * lv[0][0, 0, 0])
    assert np.allclose(oracle.sample_lod(lv, 1, [-0.5, 0.5, 0.5], 0.0), 0.0)
    p = [0.3, 0.6, 0.7]
    assert np.allclose(oracle.sample_lod(lv, 5, p, 9.0), oracle.sample_lod(lv, 5, p, 3.0))
    assert np.allclose(oracle.sample_lod(lv, 5, p, -2.0), oracle.sample_lod(lv, 5, p, 0.0))


# --------------------------------------------------------------------------------------
# trace_cone / trace_diffuse (voxel_cone_tracing.frag:88-168)
# --------------------------------------------------------------------------------------


def _empty_levels(res, levels, dt=np.float32, value=None):
    out = []
    for l in range(levels):
        s = res >> l
        a = np.zeros((s, s, s, 4), dt) if value is None else np.broadcast_to(np.asarray(value, dt),
                                                                              (s, s, s, 4)).copy()
        out.append(a if l == 0 else [a] * 6)
    return out


@pytest.mark.parametrize("res,steps", [(64, 15), (128, 18), (256, 21), (512, 23)])
def test_diffuse_cone_step_count_closed_form(res, steps):
    # In an empty volume the march only stops at sqrt(3): the step sequence is
    # fixed by the aperture (0.55785173935) and the voxel size alone.
    lv = _empty_levels(res, 7)
    out, n = oracle.trace_cone(lv, [0.5, 0.5, 0.5], [0.0, 0.0, 1.0], 0.55785173935, 1.73205080757)
    assert n == steps
    assert np.all(out == 0)


def test_opaque_uniform_volume_stops_after_first_sample():
    v = [0.2, 0.4, 0.6, 1.0]
    lv = _empty_levels(32, 6, value=v)
    d = np.array([1.0, 2.0, 2.0]) / 3.0
    out, n = oracle.trace_cone(lv, [0.5, 0.5, 0.5], d, 0.55785173935, 1.73205080757)
    # sample_voxel weights are |d| unnormalised: |dx| + |dy| + |dz| = 5/3
    assert n == 1
    assert np.allclose(out, np.array(v) * 5.0 / 3.0, rtol=1e-5)


def test_diffuse_directions_for_canonical_normals():
    # n = +z: tangent(n) = normalize(n x y) = (-1, 0, 0); o2 = o1 x n = (0, 1, 0)
    d = oracle.diffuse_dirs([0.0, 0.0, 1.0])
    assert np.allclose(d[0], [0, 0, 1])
    assert np.allclose(d[1], [-0.5, 0, 0.5])
    assert np.allclose(d[3], [0, 0.5, 0.5])
    assert np.allclose(d[5], [-0.25, 0.25, 0.5])
    # n = +x: |n x z| = |n x y| = 1, the tie goes to n x y = (0, 0, 1)
    d = oracle.diffuse_dirs([1.0, 0.0, 0.0])
    assert np.allclose(d[1], [0.5, 0, 0.5])
    # n = +y: n x z = (1, 0, 0) is longer than n x y = 0
    d = oracle.diffuse_dirs([0.0, 1.0, 0.0])
    assert np.allclose(d[1], [0.5, 0.5, 0])
    # every cone leans 45 degrees or less from the normal, all on its side
    rng = np.random.default_rng(2)
    for _ in range(50):
        n = rng.standard_normal(3)
        n /= np.linalg.norm(n)
        dd = oracle.diffuse_dirs(n)
        cosines = (dd @ n) / np.linalg.norm(dd, axis=1)
        assert np.all(cosines >= math.cos(math.radians(45)) - 1e-9)


# --------------------------------------------------------------------------------------
# rasteriser (GL pixel centres, 1/256 snap, top-left rule) and dominant axis
# --------------------------------------------------------------------------------------


def test_raster_is_watertight_on_shared_edges():
    rng = np.random.default_rng(6)
    W = H = 48
    for _ in range(200):
        # a random convex quad (points on an ellipse, counter-clockwise) split
        # along a diagonal, plus a fan around its centre
        ang = np.sort(rng.random(4) * 2 * np.pi)
        ax, ay = rng.random(2) * 16 + 4
        cx, cy = rng.random(2) * 8 + 20
        p = np.stack([cx + ax * np.cos(ang), cy + ay * np.sin(ang)], 1)
        c = p.mean(0)
        m1 = oracle.raster_coverage(p[[0, 1, 2], 0], p[[0, 1, 2], 1], W, H)
        m2 = oracle.raster_coverage(p[[0, 2, 3], 0], p[[0, 2, 3], 1], W, H)
        assert (m1.astype(int) + m2).max() <= 1  # no pixel twice
        fan = sum(oracle.raster_coverage([c[0], p[k, 0], p[(k + 1) % 4, 0]], [c[1], p[k, 1], p[(k + 1) % 4, 1]],
                                         W, H).astype(int) for k in range(4))
        assert fan.max() <= 1
        assert np.array_equal(fan, (m1 + m2).astype(int))  # same pixels, no gaps


def test_raster_pixel_centres_on_grid_aligned_edges():
    # axis-aligned right triangle with edges exactly through pixel centres:
    # the top-left rule keeps the left edge and the bottom... only once.
    m = oracle.raster_coverage([2.5, 10.5, 2.5], [2.5, 2.5, 10.5], 16, 16)
    # centres (i + 0.5, j + 0.5) with i, j >= 2 and i + j <= 10 - strictly or on edges
    cnt = int(m.sum())
    interior = sum(1 for i in range(16) for j in range(16) if i > 2 and j > 2 and i + j < 10)
    assert interior < cnt <= interior + 8 + 8 + 9
    # reversed winding covers exactly the same pixels
    m2 = oracle.raster_coverage([2.5, 2.5, 10.5], [2.5, 10.5, 2.5], 16, 16)
    assert np.array_equal(m, m2)


def test_raster_area_matches_for_large_triangles():
    m = oracle.raster_coverage([3.3, 900.7, 100.1], [5.2, 40.9, 880.4], 1024, 1024)
    area = 0.5 * abs((900.7 - 3.3) * (880.4 - 5.2) - (40.9 - 5.2) * (100.1 - 3.3))
    perim = sum(math.dist(a, b) for a, b in [((3.3, 5.2), (900.7, 40.9)), ((900.7, 40.9), (100.1, 880.4)),
                                               ((100.1, 880.4), (3.3, 5.2))])
    assert abs(int(m.sum()) - area) <= perim


def test_dominant_axis_and_ties_go_to_xz():
    assert oracle.dominant_axis([0, 0, 0, 1, 0, 0, 0, 1, 0]) == 2  # in the XY plane
    assert oracle.dominant_axis([0, 0, 0, 0, 1, 0, 0, 0, 1]) == 0  # YZ
    assert oracle.dominant_axis([0, 0, 0, 1, 0, 0, 0, 0, 1]) == 1  # XZ
    # |n| equal in x and z: neither strict test passes -> XZ
    assert oracle.dominant_axis([0, 0, 0, 0, 1, 0, 1, 0, -1]) == 1
    # equal in y and z: z > y fails, x > y fails -> XZ
    assert oracle.dominant_axis([0, 0, 0, 1, 0, 0, 0, 1, -1]) == 1
    # equal in x and y: z is smaller, x > y fails -> XZ
    assert oracle.dominant_axis([0, 0, 0, 0, 0, 1, 1, -1, 0]) == 1


# --------------------------------------------------------------------------------------
# several draws in one frame (renderer.cpp:242-256: each queued model is one draw)
# --------------------------------------------------------------------------------------


def _soup_draws(n_draws, model_at=()):
    from helpers import SceneSpec, model_matrix

    sc = SceneSpec("soup", res=32, ntri=3000)
    cuts = np.linspace(0, 3000, n_draws + 1).astype(int)
    draws = [(model_matrix("rot_scale") if k in model_at else sc.M, sc.pos[a:b], sc.nrm[a:b], sc.tri_mat[a:b])
             for k, (a, b) in enumerate(zip(cuts[:-1], cuts[1:]))]
    return sc, draws


def test_voxelize_split_into_draws_is_the_single_draw():
    """Identity matrices: the draws' triangles in submission order are the
    single draw's triangles in index order, so every output, the
    order-dependent running average included, is the same."""
    sc, draws = _soup_draws(7)
    one = oracle.voxelize(sc.res, sc.cube_size, sc.M, sc.pos, sc.nrm, sc.tri_mat, sc.mat_records, sc.lights)
    many = oracle.voxelize_draws(sc.res, sc.cube_size, draws, sc.mat_records, sc.lights)
    assert one["fragments"] > 300
    for k in ("ref", "q32", "count"):
        assert np.array_equal(one[k], many[k]), k


def test_voxelize_draws_each_take_their_own_matrix():
    """Fragments of a draw depend only on that draw's matrix: counts add up
    over the draws, and so do the double-precision colour sums; a draw
    voxelized with another draw's matrix moves fragments and fails both."""
    sc, draws = _soup_draws(3, model_at=(1,))
    both = oracle.voxelize_draws(sc.res, sc.cube_size, draws, sc.mat_records, sc.lights)
    parts = [oracle.voxelize_draws(sc.res, sc.cube_size, [d], sc.mat_records, sc.lights) for d in draws]
    assert np.array_equal(both["count"], sum(p["count"] for p in parts))
    s_both = both["f64"] * both["count"][..., None]
    s_parts = sum(p["f64"] * p["count"][..., None] for p in parts)
    assert np.allclose(s_both, s_parts, rtol=1e-12, atol=1e-12)
    # the transformed draw lands elsewhere than it would untransformed
    plain = oracle.voxelize_draws(sc.res, sc.cube_size, [(sc.M,) + draws[1][1:]], sc.mat_records, sc.lights)
    assert not np.array_equal(plain["count"], parts[1]["count"])
