"""Pins of the oracle's multi-bounce bake (orc_bake_indirect), CPU.

The indirect term adds albedo x scale x the probe-grid lookup at each
Lambert hit.  Checked against what the mathematics fixes, not against a
retyped formula:
- zero albedo reproduces orc_bake bit for bit;
- a constant source grid (c00 = C) inside a closed box: every ray hits, the
  lookup returns C Y00 anywhere (weights sum to one), so the result is
  albedo x scale x C x Y00 x (the rays' quadrature of Y_k) -- c00 gains
  exactly albedo x scale x C;
- an l = 1 source field L(d) = d_y over a floor whose triangles face away
  from the probes: the clamped-cosine convolution gives irradiance/pi =
  (2/3) n_y, and the two-sided normal facing the downward rays is +y, so
  each hit carries albedo x (2/3); a missing flip gives -2/3, a missing band
  factor 1.
"""
import numpy as np
import pytest

import orc
import vlb

Y00 = 0.28209479177387814
Y1 = 0.48860251190291992


def _box(h=2.0):
    v = np.array([[x, y, z] for x in (-h, h) for y in (-h, h) for z in (-h, h)], np.float32)
    q = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    idx = np.array([t for a, b, c, d in q for t in ((a, b, c), (a, c, d))], np.uint32)
    return v, idx


def _grid(lo, hi, cnt):
    probes, step = vlb.probe_grid(lo, hi, cnt)
    return probes, vlb.grid_desc(lo, step, cnt, num_coefs=9)


def _params(probes, rays=512, **kw):
    base = dict(probes=probes, rays=rays, sh_order=2, shade_mode=vlb.SHADE_LAMBERT, lights=[],
                sky_zenith=(0.0, 0.0, 0.0), sky_horizon=(0.0, 0.0, 0.0), sun_rgb=(0.0, 0.0, 0.0),
                t_max=1e7)
    base.update(kw)
    return vlb.Params(**base)


def _quad(R, mask=None):
    d = orc.fibonacci(R).astype(np.float64)
    Y = orc.sh_eval(d[:, [0, 2, 1]], K=9)          # the bake's SH frame (x, z, y)
    w = np.full(R, 4.0 * np.pi / R)
    if mask is not None:
        w = w * mask
    return (w[:, None] * Y).sum(axis=0)              # sum_r w_r Y_k(d_r)


@pytest.mark.parametrize("mode64", [True, False])
def test_zero_albedo_is_the_direct_bake(mode64):
    s = vlb.generate_scene(vlb.SCENE_ROOM, 2)
    alb = np.zeros((s["idx"].shape[0], 3), np.float32)
    sc = orc.Scene(s["xyz"], s["idx"], normals=s.get("normals"), albedo=alb)
    probes, g = _grid([2.0, 1.0, 2.0], [8.0, 3.0, 8.0], [2, 2, 2])
    sun = vlb.sun_direction(np.random.default_rng(1))
    p = _params(probes, 256, lights=[(vlb.LIGHT_DIRECTIONAL, (-sun).tolist(), (3.0, 3.0, 3.0))],
                sky_zenith=(2.0, 1.8, 1.5), sky_horizon=(0.5, 0.6, 0.7), sun_rgb=(50.0, 45.0, 40.0))
    src = np.random.default_rng(2).normal(size=(8, 9, 3)).astype(np.float32)
    for vis in (False, True):
        lp = vlb.lookup_params(vlb.LOOKUP_IRRADIANCE, vis, 1e-3)
        a = sc.bake_indirect(p, g, src, lp, 1.0, mode64=mode64)["coeffs"]
        b = sc.bake(p, mode64=mode64)["coeffs"]
        assert np.array_equal(a, b)


@pytest.mark.parametrize("vis", [False, True])
@pytest.mark.parametrize("scale", [1.0, 0.25])
def test_constant_grid_in_closed_box(vis, scale):
    v, idx = _box()
    alb = np.tile(np.array([[0.8, 0.5, 0.2]], np.float32), (12, 1))
    sc = orc.Scene(v, idx, albedo=alb)
    probes, g = _grid([-0.5, -0.5, -0.5], [0.5, 0.5, 0.5], [2, 2, 2])
    R = 512
    p = _params(probes, R)
    C = np.array([3.0, 2.0, 1.0])
    src = np.zeros((8, 9, 3), np.float32)
    src[:, 0, :] = C
    lp = vlb.lookup_params(vlb.LOOKUP_IRRADIANCE, vis, 1e-3)
    got = sc.bake_indirect(p, g, src, lp, scale)["coeffs"]
    a = alb[0].astype(np.float64)
    want = a * scale * C * Y00 * _quad(R)[:, None]       # (9, 3)
    np.testing.assert_allclose(got, np.broadcast_to(want, got.shape), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(got[:, 0, :], np.broadcast_to(a * scale * C, (8, 3)), rtol=1e-9)


def test_first_order_field_over_a_floor_facing_away():
    L = 1e5
    v = np.array([[-L, 0, -L], [L, 0, -L], [L, 0, L], [-L, 0, L]], np.float32)
    idx = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)     # geometric normal -y
    alb = np.array([[0.6, 0.6, 0.6]] * 2, np.float32)
    sc = orc.Scene(v, idx, albedo=alb)
    probes, g = _grid([-1.0, 1.0, -1.0], [1.0, 2.0, 1.0], [2, 2, 2])
    R = 1024
    p = _params(probes, R)
    f = np.array([1.0, 0.5, 0.25])
    src = np.zeros((8, 9, 3), np.float32)
    src[:, 2, :] = f / Y1            # L(d) = f d_y: SH index 2 is Y_1^0 ~ world y
    lp = vlb.lookup_params(vlb.LOOKUP_IRRADIANCE, False, 1e-3)
    r = sc.bake_indirect(p, g, src, lp, 1.0, want_hits=True)
    hits = r["hits"] >= 0
    d = orc.fibonacci(R)
    assert np.array_equal(hits, np.broadcast_to(d[:, 1] < 0, hits.shape))  # every downward ray hits
    for i in range(8):
        f_eff = src[0, 2, :].astype(np.float64) * Y1     # f after float32 storage
        want = float(np.float32(0.6)) * (2.0 / 3.0) * f_eff[None, :] * _quad(R, hits[i].astype(np.float64))[:, None]
        np.testing.assert_allclose(r["coeffs"][i], want, rtol=1e-9, atol=1e-12)


def test_rejects_reference_shading_and_mismatched_grid():
    v, idx = _box()
    sc = orc.Scene(v, idx)
    probes, g = _grid([-0.5, -0.5, -0.5], [0.5, 0.5, 0.5], [2, 2, 2])
    src = np.zeros((8, 9, 3), np.float32)
    lp = vlb.lookup_params(vlb.LOOKUP_IRRADIANCE, False, 1e-3)
    with pytest.raises(ValueError):
        sc.bake_indirect(_params(probes, shade_mode=vlb.SHADE_REFERENCE), g, src, lp)
    with pytest.raises(ValueError):      # K of the grid differs from the bake's
        sc.bake_indirect(_params(probes, sh_order=3), g, src, lp)
