"""GPU parity of the probe-grid lookup (vlb_probe_lookup) against the oracle.

A grid is baked on the GPU, surface points come from rays traced into the
room, and the lookup's value per point is compared with orc_probe_lookup on
the same coefficients: |got - ref| <= 1e-5 x max(|ref| of the point, median
|ref|) -- FP32 against double, where the un-convolved radiance at a normal
can nearly cancel; a probe whose visibility differed moves a point's value by
its weight share times the probes' difference, far more than that.
"""
import numpy as np
import pytest

import orc
import vlb

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def baked(ctx):
    s = vlb.generate_scene(vlb.SCENE_ROOM, 2)
    sid = ctx.upload_scene_dict(s)
    out = {}
    for order in (vlb.ORDER_LINEAR, vlb.ORDER_REFERENCE_BFS):
        b0, b1, cnt = [0.3125, 0.25, 0.3125], [9.6875, 3.75, 9.6875], [8, 4, 8]
        probes, step = vlb.probe_grid(b0, b1, cnt, order)
        rng = np.random.default_rng(7)
        sun = vlb.sun_direction(rng)
        p = vlb.Params(probes=probes, rays=1024, sh_order=2,
                       lights=[(vlb.LIGHT_DIRECTIONAL, (-sun).tolist(), (3.0, 2.5, 2.0))],
                       sun_dir=sun.tolist(), shadow_bias=1e-3)
        out[order] = (ctx.bake(sid, p), vlb.grid_desc(b0, step, cnt, order=order, num_coefs=9))
    yield s, sid, orc.Scene.from_dict(s), out
    ctx.release_scene(sid)


def _surface_points(ctx, sid, s, n, seed):
    rng = np.random.default_rng(seed)
    org = rng.uniform([0.5, 0.5, 0.5], [9.5, 3.5, 9.5], size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    prim, t, _, _ = ctx.trace_closest(sid, org, d, 0.001, 1e4)
    hit = prim >= 0
    pts = (org + d * t[:, None])[hit]
    tri = s["xyz"][s["idx"][prim[hit]]]
    nrm = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    flip = (nrm * d[hit]).sum(axis=1) > 0          # face the incoming ray
    nrm[flip] *= -1
    return pts.astype(np.float32), nrm.astype(np.float32)


@pytest.mark.parametrize("order", [vlb.ORDER_LINEAR, vlb.ORDER_REFERENCE_BFS])
@pytest.mark.parametrize("mode", [vlb.LOOKUP_REFERENCE, vlb.LOOKUP_IRRADIANCE])
@pytest.mark.parametrize("vis", [False, True])
def test_lookup_matches_oracle(ctx, baked, order, mode, vis):
    s, sid, osc, grids = baked
    coeffs, g = grids[order]
    pts, nrm = _surface_points(ctx, sid, s, 4000, 11)
    lp = vlb.lookup_params(mode, vis, 2e-3)
    got = ctx.probe_lookup(sid, g, coeffs, lp, pts, nrm)
    ref = orc.probe_lookup(osc, g, coeffs, lp, pts, nrm)
    mag = np.abs(ref).max(axis=1)
    err = np.abs(got - ref).max(axis=1) / np.maximum(mag, np.median(mag))
    i = int(np.argmax(err))
    assert err.max() <= 1e-5, f"max rel err {err.max():.3e} at {pts[i]} n={nrm[i]} got={got[i]} ref={ref[i]}"
    if vis:  # the walls and boxes do hide probes from some points
        lp0 = vlb.lookup_params(mode, False, 2e-3)
        assert not np.allclose(got, ctx.probe_lookup(sid, g, coeffs, lp0, pts, nrm))


def test_lookup_l3_literal_constants(ctx, baked):
    s, sid, osc, grids = baked
    coeffs9, g9 = grids[vlb.ORDER_LINEAR]
    rng = np.random.default_rng(3)
    coeffs = rng.normal(size=(coeffs9.shape[0], 16, 3)).astype(np.float32)
    g = vlb.grid_desc(g9.origin[:], g9.step[:], g9.count[:], num_coefs=16, sh_constants=vlb.SH_LITERAL)
    pts, nrm = _surface_points(ctx, sid, s, 1000, 12)
    for mode in (vlb.LOOKUP_REFERENCE, vlb.LOOKUP_IRRADIANCE):
        lp = vlb.lookup_params(mode, True, 2e-3)
        got = ctx.probe_lookup(sid, g, coeffs, lp, pts, nrm)
        ref = orc.probe_lookup(osc, g, coeffs, lp, pts, nrm)
        np.testing.assert_allclose(got, ref, rtol=1e-4, atol=1e-5)
