"""GPU parity: libvlb.so (HIP, gfx950) against the CPU oracle on the same inputs.

Bars: hit prim ids, t, u, v and shadow visibility are bit-exact (integer /
discrete results of the same IEEE operation sequence); SH coefficients of the
FP32 path are within 1e-5 relative L2 of the FP64 oracle, BF16/FP16 within
1e-3 (rel. L2 per probe and channel).
"""
import numpy as np
import pytest

import orc
import vlb
from conftest import rel_err

pytestmark = pytest.mark.gpu


def _room(seed=2):
    return vlb.generate_scene(vlb.SCENE_ROOM, seed)


def _rays(n, seed, lo=(0.2, 0.2, 0.2), hi=(9.8, 3.8, 9.8)):
    rng = np.random.default_rng(seed)
    org = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return org, d.astype(np.float32)


@pytest.fixture(scope="module")
def room(ctx):
    s = _room()
    sid = ctx.upload_scene_dict(s)
    yield s, sid, orc.Scene.from_dict(s)
    ctx.release_scene(sid)


def test_trace_closest_bit_exact(ctx, room):
    s, sid, o = room
    org, d = _rays(20000, 1)
    gp, gt, gu, gv = ctx.trace_closest(sid, org, d, 0.001, 1e4)
    cp, ct, cu, cv = o.trace_closest(org, d, 0.001, 1e4)
    assert (gp == cp).all(), f"prim mismatch on {(gp != cp).sum()} rays"
    hit = cp >= 0
    assert hit.mean() > 0.5
    assert np.array_equal(gt[hit].view(np.uint32), ct[hit].view(np.uint32))
    assert np.array_equal(gu[hit].view(np.uint32), cu[hit].view(np.uint32))
    assert np.array_equal(gv[hit].view(np.uint32), cv[hit].view(np.uint32))


def test_trace_occluded_bit_exact(ctx, room):
    s, sid, o = room
    org, d = _rays(20000, 2)
    tmax = np.random.default_rng(3).uniform(0.05, 12.0, size=20000).astype(np.float32)
    g = ctx.trace_occluded(sid, org, d, tmax, 0.0)
    c = o.trace_occluded(org, d, tmax, 0.0)
    assert (g == c).all(), f"visibility mismatch on {(g != c).sum()} rays"
    assert 0.05 < c.mean() < 0.95


def test_bvh_export_invariants(ctx, room):
    s, sid, _ = room
    nodes, leaf = ctx.export_bvh(sid)
    nt = s["idx"].shape[0]
    assert sorted(leaf.tolist()) == list(range(nt))
    tri = s["xyz"][s["idx"]]  # (nt, 3, 3)
    lo_t, hi_t = tri.min(axis=1), tri.max(axis=1)
    ch = nodes[:, 12:14].view(np.int32)
    # every leaf's box in its parent equals the exact triangle bounds
    for slot, (a, b) in enumerate(((0, 6), (6, 12))):
        box = nodes[:, a:b]
        lo = box[:, [0, 1, 2]] if slot == 0 else box[:, [0, 1, 2]]
        c = ch[:, slot]
        m = c < 0
        prim = leaf[~c[m]]
        blo = np.stack([nodes[m, a], nodes[m, a + 1], nodes[m, a + 2]], 1)
        bhi = np.stack([nodes[m, a + 3], nodes[m, a + 4], nodes[m, a + 5]], 1)
        assert np.array_equal(blo, lo_t[prim]) and np.array_equal(bhi, hi_t[prim])
    # inner children referenced exactly once, leaves exactly once
    inner = ch[ch >= 0]
    assert len(np.unique(inner)) == len(inner) == nodes.shape[0] - 1
    assert (ch < 0).sum() == nt
    # an inner child's box in its parent is the union of its own two child
    # boxes; the root's two boxes together cover the whole scene
    for a, slot in ((0, 0), (6, 1)):
        c = ch[:, slot]
        m = c >= 0
        kid = nodes[c[m]]
        ulo = np.minimum(kid[:, 0:3], kid[:, 6:9])
        uhi = np.maximum(kid[:, 3:6], kid[:, 9:12])
        assert np.array_equal(nodes[m, a:a + 3], ulo) and np.array_equal(nodes[m, a + 3:a + 6], uhi)
    root_lo = np.minimum(nodes[0, 0:3], nodes[0, 6:9])
    root_hi = np.maximum(nodes[0, 3:6], nodes[0, 9:12])
    assert np.array_equal(root_lo, lo_t.min(axis=0)) and np.array_equal(root_hi, hi_t.max(axis=0))
    assert ctx.scene_info(sid)["depth"] <= 64


def test_bvh_coincident_and_degenerate_triangles(ctx):
    """Thousands of copies of one triangle, and of a zero-area one, among a
    few random ones: the build stays shallow (equal-area neighbours pair up
    instead of chaining), and closest hits still go to the lowest prim id."""
    rng = np.random.default_rng(9)
    one = np.array([[0, 0, 0], [4, 0, 0], [0, 4, 0]], np.float32)
    point = np.full((3, 3), 2.0, np.float32)
    rand = rng.uniform(-6, 6, size=(300, 3, 3)).astype(np.float32)
    tris = np.concatenate([np.repeat(one[None], 3000, 0), np.repeat(point[None], 3000, 0), rand])
    tris = tris[rng.permutation(len(tris))]
    xyz = tris.reshape(-1, 3)
    idx = np.arange(len(xyz), dtype=np.uint32).reshape(-1, 3)
    sid = ctx.upload_scene(xyz, idx)
    assert ctx.scene_info(sid)["depth"] <= 48
    o = orc.Scene(xyz, idx)
    org = rng.uniform([-1, -1, -5], [5, 5, -1], size=(5000, 3)).astype(np.float32)
    tgt = rng.uniform([-1, -1, -1], [5, 5, 1], size=(5000, 3)).astype(np.float32)
    d = tgt - org
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    gp, gt, _, _ = ctx.trace_closest(sid, org, d, 0.0, 1e4)
    cp, ct, _, _ = o.trace_closest(org, d, 0.0, 1e4)
    assert (gp == cp).all(), f"prim mismatch on {(gp != cp).sum()} rays"
    assert np.array_equal(gt.view(np.uint32), ct.view(np.uint32))
    assert (cp >= 0).mean() > 0.3
    ctx.release_scene(sid)


def test_deep_tree_bakes_identical_across_kernel_families(ctx):
    """A deep BVH (thousands of coincident and degenerate triangles among
    random ones) makes traversal stacks run past the LDS entries of every
    kernel family (16 in the fused kernels, 12 in the wavefront ones) into
    their global spill areas: every family bakes the same bits, and the
    oracle's coefficients."""
    rng = np.random.default_rng(19)
    one = np.array([[0, 0, 0], [4, 0, 0], [0, 4, 0]], np.float32)
    point = np.full((3, 3), 2.0, np.float32)
    rand = rng.uniform(-6, 6, size=(400, 3, 3)).astype(np.float32)
    tris = np.concatenate([np.repeat(one[None], 3000, 0), np.repeat(point[None], 3000, 0), rand])
    tris = tris[rng.permutation(len(tris))]
    xyz = tris.reshape(-1, 3)
    idx = np.arange(len(xyz), dtype=np.uint32).reshape(-1, 3)
    sid = ctx.upload_scene(xyz, idx)
    # stack bound 3 * ceil(depth / 2) + 1 >= 19 entries: past both LDS depths
    assert ctx.scene_info(sid)["depth"] >= 12
    o = orc.Scene(xyz, idx)
    probes = rng.uniform([-1, -1, -2], [5, 5, 2], size=(24, 3)).astype(np.float32)
    p = _bake_params(probes, rays=400)
    ref = None
    for v in (12, 36, 16, 20, 1):
        p.trace_variant = v
        g = ctx.bake(sid, p)
        if ref is None:
            ref = g
            assert rel_err(g, o.bake(p, mode64=True)["coeffs"]).max() <= 1e-5
        assert np.array_equal(g.view(np.uint32), ref.view(np.uint32)), v
    ctx.release_scene(sid)


def _bake_params(probes, rays=1024, **kw):
    rng = np.random.default_rng(7)
    sun = vlb.sun_direction(rng)
    base = dict(probes=probes, rays=rays, sh_order=2, shade_mode=vlb.SHADE_LAMBERT,
                lights=[(vlb.LIGHT_DIRECTIONAL, (-sun).tolist(), (3.0, 2.5, 2.0))],
                sun_dir=sun.tolist(), sky_zenith=(2.0, 1.8, 1.5), sky_horizon=(0.5, 0.6, 0.7),
                sun_rgb=(50.0, 45.0, 40.0), shadow_bias=1e-3)
    base.update(kw)
    return vlb.Params(**base)


def _probes(n, seed=5):
    rng = np.random.default_rng(seed)
    return rng.uniform([0.3, 0.3, 0.3], [9.7, 3.7, 9.7], size=(n, 3)).astype(np.float32)


def test_bake_fp32_vs_fp64_oracle(ctx, room):
    s, sid, o = room
    p = _bake_params(_probes(48), rays=1024)
    g = ctx.bake(sid, p)
    ref = o.bake(p, mode64=True)["coeffs"]
    err = rel_err(g, ref)
    assert err.max() <= 1e-5, f"max rel err {err.max():.3e}"


def test_bake_radiance_and_hits_match_oracle(ctx, room):
    s, sid, o = room
    p = _bake_params(_probes(20, seed=9), rays=256)
    rad, hits = ctx.bake_radiance(sid, p)
    ref = o.bake(p, mode64=False, want_radiance=True, want_hits=True)
    assert np.array_equal(hits, ref["hits"]), f"{(hits != ref['hits']).sum()} hit ids differ"
    np.testing.assert_allclose(rad, ref["radiance"], rtol=2e-6, atol=1e-7)


@pytest.mark.parametrize("prec,bound", [(vlb.PREC_BF16, 1e-4), (vlb.PREC_FP16, 1e-4),
                                        (vlb.PREC_BF16_FAST, 2e-2), (vlb.PREC_FP16_FAST, 3e-3)])
def test_bake_16bit_projection_error(ctx, room, prec, bound):
    # split (hi/lo) 16-bit forms must meet the FP32 bar within 1e-4; the
    # single-product forms carry the 16-bit rounding of Y00 and of flat-lit
    # radiance runs (bounded, not averaged away)
    s, sid, o = room
    p = _bake_params(_probes(48), rays=2048, precision=prec)
    g = ctx.bake(sid, p)
    p64 = _bake_params(_probes(48), rays=2048)
    ref = o.bake(p64, mode64=True)["coeffs"]
    err = rel_err(g, ref)
    assert err.max() <= bound, f"max rel err {err.max():.3e}"


def test_bake_l3_exact_and_literal(ctx, room):
    s, sid, o = room
    for const in (vlb.SH_EXACT, vlb.SH_LITERAL):
        p = _bake_params(_probes(16), rays=512, sh_order=3, sh_constants=const)
        g = ctx.bake(sid, p)
        ref = o.bake(p, mode64=True)["coeffs"]
        assert g.shape == (16, 16, 3)
        assert rel_err(g, ref).max() <= 1e-5


def test_bake_reference_shading(ctx, room):
    s, sid, o = room
    for flags in (0, vlb.REF_SRGB, vlb.REF_SRGB | vlb.REF_UNORM8):
        p = _bake_params(_probes(16, seed=11), rays=512, shade_mode=vlb.SHADE_REFERENCE,
                         ref_light=(5.0, 3.5, 5.0), ref_flags=flags, sh_order=3)
        g = ctx.bake(sid, p)
        ref = o.bake(p, mode64=True)["coeffs"]
        assert rel_err(g, ref).max() <= 1e-5, flags


def test_bake_point_lights(ctx, room):
    s, sid, o = room
    p = _bake_params(_probes(32, seed=13), rays=512,
                     lights=[(vlb.LIGHT_POINT, (5.0, 3.5, 5.0), (80.0, 60.0, 40.0)),
                             (vlb.LIGHT_POINT, (2.0, 1.0, 8.0), (10.0, 30.0, 50.0)),
                             (vlb.LIGHT_DIRECTIONAL, (0.3, -0.9, 0.3), (1.0, 1.0, 1.0))])
    g = ctx.bake(sid, p)
    ref = o.bake(p, mode64=True)["coeffs"]
    assert rel_err(g, ref).max() <= 1e-5


@pytest.mark.parametrize("P,R", [(1, 1), (1, 33), (17, 31), (33, 100)])
def test_bake_ragged_sizes(ctx, room, P, R):
    s, sid, o = room
    p = _bake_params(_probes(P, seed=P + R), rays=R)
    g = ctx.bake(sid, p)
    ref = o.bake(p, mode64=True)["coeffs"]
    assert rel_err(g, ref).max() <= 1e-5


def test_bake_equirect_reference_dirs(ctx, room):
    s, sid, o = room
    p = _bake_params(_probes(4, seed=3), dir_mode=vlb.DIRS_EQUIRECT, equirect_w=63,
                     equirect_h=20, sh_order=3, shade_mode=vlb.SHADE_REFERENCE)
    g = ctx.bake(sid, p)
    ref = o.bake(p, mode64=True)["coeffs"]
    assert rel_err(g, ref).max() <= 1e-5


def test_empty_and_single_triangle_scenes(ctx):
    p = _bake_params(_probes(5), rays=300)
    sid = ctx.upload_scene(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32))
    g = ctx.bake(sid, p)
    ref = orc.Scene(np.zeros((0, 3)), np.zeros((0, 3), np.uint32)).bake(p)["coeffs"]
    assert rel_err(g, ref).max() <= 1e-5
    ctx.release_scene(sid)
    xyz = np.array([[0, 2, 0], [10, 2, 0], [0, 2, 10]], np.float32)
    idx = np.array([[0, 1, 2]], np.uint32)
    sid = ctx.upload_scene(xyz, idx, albedo=[[0.5, 0.6, 0.7]])
    g = ctx.bake(sid, p)
    ref = orc.Scene(xyz, idx, albedo=[[0.5, 0.6, 0.7]]).bake(p)["coeffs"]
    assert rel_err(g, ref).max() <= 1e-5
    ctx.release_scene(sid)


# batches of 1 (one cubemap per workgroup), 2 and 3 (two per workgroup at
# L2, the last one alone), power-of-two and other face sizes; faces up to 147
# texels read each texel's direction from a table, larger ones compute it
# (147: the largest table, 160: computed)
@pytest.mark.parametrize("N,batch", [(32, 2), (32, 1), (24, 3), (17, 5), (147, 1), (160, 2)])
def test_skybox_cubemap_vs_oracle(ctx, N, batch):
    rng = np.random.default_rng(N * batch)
    cubes = []
    for b in range(batch):
        sun = vlb.sun_direction(rng)
        cubes.append(orc.procedural_cubemap(N, sun, tint=tuple(rng.uniform(0.5, 1.0, 3))))
    cubes = np.stack(cubes)
    for order, K in ((2, 9), (3, 16)):
        g = ctx.skybox_sh(cubes, order)
        for b in range(batch):
            ref = orc.skybox_sh(cubes[b], K)
            assert rel_err(g[b], ref).max() <= 1e-5, (order, b)


@pytest.mark.parametrize("variant", [vlb.EQUIRECT_SH_COMP, vlb.EQUIRECT_SKYBOX_COMP])
def test_equirect_projection_vs_oracle(ctx, variant):
    rng = np.random.default_rng(variant)
    img = rng.integers(0, 256, size=(100, 314, 4), dtype=np.uint8)
    g = ctx.project_equirect(img, variant)
    ref = orc.equirect_project(img, variant)
    assert rel_err(g, ref).max() <= 1e-5


def test_sh_reconstruct_vs_oracle(ctx):
    rng = np.random.default_rng(4)
    c = (rng.normal(size=(16, 3)) * 2e-4).astype(np.float32)
    g = ctx.sh_reconstruct(c, 157, 50)
    ref = orc.sh_reconstruct(c, 157, 50)
    diff = np.abs(g.astype(int) - ref.astype(int))
    assert diff.max() <= 1 and (diff > 0).mean() < 1e-3


def test_plan_graph_replays_match_eager(ctx, room):
    s, sid, o = room
    p = _bake_params(_probes(40), rays=512)
    pl = ctx.plan(sid, p)
    pl.execute()
    eager = pl.download()
    pl.use_graph(True)
    for _ in range(3):
        pl.execute()
        got = pl.download()
        assert np.array_equal(got, eager)
        ctx.sync()
    pl.destroy()


# ---- every trace pipeline computes the same bits: fused k_trace (1..12), the
# wavefront k_rays/k_shade pipeline (13..24) and fused with far-first shadow
# queries (25..36)
VARIANTS = [1, 3, 4, 8, 12, 13, 15, 16, 20, 23, 25, 28, 32, 36]


def test_trace_variants_are_bitwise_identical(ctx, room):
    s, sid, o = room
    p = _bake_params(_probes(40, seed=21), rays=700)
    ref = None
    for v in VARIANTS:
        p.trace_variant = v
        g = ctx.bake(sid, p)
        if ref is None:
            ref = g
            assert rel_err(g, o.bake(p, mode64=True)["coeffs"]).max() <= 1e-5
        assert np.array_equal(g.view(np.uint32), ref.view(np.uint32)), v


@pytest.mark.parametrize("mode", [vlb.SHADE_LAMBERT, vlb.SHADE_REFERENCE])
def test_far_first_shadows_same_bits_and_shadow_rays(ctx, room, mode):
    """Shadow queries visiting children far first (25..36) answer as the
    slot-order ones (1..12): the same coefficients bit for bit and the same
    shadow rays, with a different node-visit count from the counting pass."""
    s, sid, o = room
    p = _bake_params(_probes(33, seed=5), rays=500, shade_mode=mode, ref_light=(5.0, 3.5, 5.0))
    got = {}
    for v in (12, 36, 25):
        p.trace_variant = v
        pl = ctx.plan(sid, p)
        pl.execute()
        g = pl.download()
        pl.count()
        got[v] = (g, pl.stats())
        pl.destroy()
    assert rel_err(got[12][0], o.bake(p, mode64=True)["coeffs"]).max() <= 1e-5
    for v in (36, 25):
        assert np.array_equal(got[v][0].view(np.uint32), got[12][0].view(np.uint32)), v
        assert got[v][1]["shadow_rays"] == got[12][1]["shadow_rays"] > 0
        assert got[v][1]["primary_rays"] == got[12][1]["primary_rays"]
    assert got[36][1]["node_visits"] == got[25][1]["node_visits"]
    assert got[36][1]["node_visits"] != got[12][1]["node_visits"]


def test_trace_variant_range(ctx, room):
    s, sid, o = room
    for v in (37, 100):
        with pytest.raises(vlb.VlbError):
            ctx.bake(sid, _bake_params(_probes(4), rays=64, trace_variant=v))


@pytest.mark.parametrize("variant", [15, 16])
def test_wavefront_radiance_and_hits(ctx, room, variant):
    s, sid, o = room
    p = _bake_params(_probes(20, seed=9), rays=256, trace_variant=variant)
    rad, hits = ctx.bake_radiance(sid, p)
    ref = o.bake(p, mode64=False, want_radiance=True, want_hits=True)
    assert np.array_equal(hits, ref["hits"])
    np.testing.assert_allclose(rad, ref["radiance"], rtol=2e-6, atol=1e-7)


# scale and offset of the room: tiny, far from the origin, large.  The
# wavefront kernels read the quantized BVH4 (8-bit child boxes on per-node
# grids), the fused ones the float BVH4; both must give the oracle's hits.
@pytest.mark.parametrize("scale,offset", [(1e-3, (0.0, 0.0, 0.0)), (1.0, (1e4, -2e4, 3e4)),
                                          (50.0, (0.0, 0.0, 0.0))])
def test_quantized_and_float_bvh4_hits_at_scale(ctx, room, scale, offset):
    s, _, _ = room
    off = np.asarray(offset, np.float32)
    xyz = (np.asarray(s["xyz"], np.float32) * np.float32(scale) + off).astype(np.float32)
    sid = ctx.upload_scene(xyz, s["idx"], normals=s.get("normals"), albedo=s.get("albedo"))
    o = orc.Scene(xyz, s["idx"], normals=s.get("normals"), albedo=s.get("albedo"))
    probes = (_probes(24, seed=13) * np.float32(scale) + off).astype(np.float32)
    for v in (12, 16):
        p = _bake_params(probes, rays=300, trace_variant=v, shadow_bias=1e-3 * scale)
        rad, hits = ctx.bake_radiance(sid, p)
        ref = o.bake(p, mode64=False, want_radiance=True, want_hits=True)
        assert np.array_equal(hits, ref["hits"]), v
        assert (hits >= 0).mean() > 0.5
        np.testing.assert_allclose(rad, ref["radiance"], rtol=2e-6, atol=1e-7)
    ctx.release_scene(sid)


def test_scene_upload_rejects_non_finite_or_huge_vertices(ctx):
    idx = np.array([[0, 1, 2]], np.uint32)
    for bad in (np.nan, np.inf, 2.0 ** 101):
        xyz = np.array([[0, 0, 0], [1, 0, 0], [0, bad, 0]], np.float32)
        with pytest.raises(vlb.VlbError):
            ctx.upload_scene(xyz, idx)


@pytest.mark.parametrize("P,R", [(1, 1), (17, 31), (33, 100)])
def test_wavefront_ragged_and_reference(ctx, room, P, R):
    s, sid, o = room
    for mode in (vlb.SHADE_LAMBERT, vlb.SHADE_REFERENCE):
        p = _bake_params(_probes(P, seed=P * R), rays=R, shade_mode=mode, ref_light=(5.0, 3.5, 5.0),
                         trace_variant=16)
        g = ctx.bake(sid, p)
        assert rel_err(g, o.bake(p, mode64=True)["coeffs"]).max() <= 1e-5


def test_wavefront_empty_scene_and_no_lights(ctx):
    p = _bake_params(_probes(5), rays=300, trace_variant=15)
    sid = ctx.upload_scene(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32))
    g = ctx.bake(sid, p)
    ref = orc.Scene(np.zeros((0, 3)), np.zeros((0, 3), np.uint32)).bake(p)["coeffs"]
    assert rel_err(g, ref).max() <= 1e-5
    ctx.release_scene(sid)


def test_wavefront_no_lights(ctx, room):
    s, sid, o = room
    p = _bake_params(_probes(8), rays=256, lights=[], trace_variant=16)
    g = ctx.bake(sid, p)
    assert rel_err(g, o.bake(p, mode64=True)["coeffs"]).max() <= 1e-5


def test_autotune_picks_a_valid_variant_and_keeps_results(ctx, room):
    s, sid, o = room
    p = _bake_params(_probes(64, seed=3), rays=1024)
    pl = ctx.plan(sid, p)
    pl.execute()
    before = pl.download()
    v = pl.autotune(1)
    assert 1 <= v <= 36
    pl.execute()
    after = pl.download()
    assert np.array_equal(before.view(np.uint32), after.view(np.uint32))
    pl.destroy()
