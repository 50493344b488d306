"""Pins of the oracle's bake (orc_bake): projection and shading, CPU.

Each case puts the probes above one large plane, where the hit point of
every ray, the normal and the light geometry have closed forms computed here
in float64 from the ray directions alone (no BVH, no shared code with the
oracle), and compares the oracle's per-ray radiance, hits and coefficients:
- sky only: coefficients = sum_r (4 pi / R) Y_k(d_r) L_r over the rays, with
  Y from the scipy-pinned basis in the bake's frame (x, z, y);
- Lambert, directional light, floor facing away from the probes: every
  downward ray carries albedo E cos / pi (two-sided normal; a missing flip
  gives zero, a dropped pi or albedo is off by a factor);
- Lambert, point light: albedo I cos / (pi d^2) per hit point;
- an occluding square between floor and light: analytic shadows;
- reference shading (env_map.rchit:73-101): albedo (ambient + Cd sDotN +
  Cs max(R.d, 0)^gloss), unflipped normal, shadow ray to the point light.
"""
import numpy as np
import pytest

import orc
import vlb

# 200 m: big enough that most downward rays hit it, small enough that the
# float32 hit point (errors grow with the vertex coordinates) stays well
# inside the 1e-3 shadow bias of the plane
L_FLOOR = 200.0


def _floor(face_up):
    v = np.array([[-L_FLOOR, 0, -L_FLOOR], [L_FLOOR, 0, -L_FLOOR], [L_FLOOR, 0, L_FLOOR],
                  [-L_FLOOR, 0, L_FLOOR]], np.float32)
    idx = np.array([[0, 2, 1], [0, 3, 2]] if face_up else [[0, 1, 2], [0, 2, 3]], np.uint32)
    return v, idx


def _probes():
    return np.array([[0.0, 1.0, 0.0], [0.7, 1.5, -0.4], [-1.2, 0.8, 2.0]], np.float32)


def _params(probes, R, **kw):
    base = dict(probes=probes, rays=R, sh_order=2, shade_mode=vlb.SHADE_LAMBERT, lights=[],
                sky_zenith=(0.0, 0.0, 0.0), sky_horizon=(0.0, 0.0, 0.0), sun_rgb=(0.0, 0.0, 0.0),
                t_min=1e-3, t_max=1e7, shadow_bias=1e-3)
    base.update(kw)
    return vlb.Params(**base)


def _floor_hits(probes, d):
    """t and hit points of every ray on y = 0 (nan where the ray goes up)."""
    o = probes.astype(np.float64)[:, None, :]
    dd = d.astype(np.float64)[None, :, :]
    with np.errstate(divide="ignore", invalid="ignore"):
        t = np.where(dd[..., 1] < 0, -o[..., 1] / dd[..., 1], np.nan)
    return t, o + dd * t[..., None]


def _project(rad, d):
    Y = orc.sh_eval(d.astype(np.float64)[:, [0, 2, 1]], K=9)          # bake's SH frame
    w = 4.0 * np.pi / d.shape[0]
    return np.einsum("rk,prc->pkc", Y * w, rad.astype(np.float64))


def test_sky_only_projection():
    sc = orc.Scene(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32))
    R = 777
    sun = np.array([0.3, 0.8, -0.52])
    sun /= np.linalg.norm(sun)
    p = _params(_probes(), R, sky_zenith=(2.0, 1.8, 1.5), sky_horizon=(0.5, 0.6, 0.7),
                sun_dir=sun.tolist(), sun_rgb=(50.0, 45.0, 40.0))
    r = sc.bake(p, mode64=True, want_radiance=True, want_hits=True)
    d = orc.fibonacci(R)
    assert (r["hits"] == -1).all()
    # the bake's projection of its own per-ray radiance (float32 output of
    # double values): sum_r w Y_k L_r
    np.testing.assert_allclose(r["coeffs"], _project(r["radiance"], d), rtol=2e-7, atol=1e-9)
    # the radiance is the configuration's sky: gradient + sun lobe
    dd = d.astype(np.float64)
    g = np.maximum(dd[:, 1], 0.0)[:, None]
    sky = (np.array([0.5, 0.6, 0.7]) + (np.array([2.0, 1.8, 1.5]) - np.array([0.5, 0.6, 0.7])) * g
           + np.array([50.0, 45.0, 40.0]) * np.exp(200.0 * (dd @ np.float32(sun).astype(np.float64) - 1.0))[:, None])
    np.testing.assert_allclose(r["radiance"][0], sky, rtol=2e-6)


@pytest.mark.parametrize("face_up", [False, True])
def test_lambert_directional_on_floor(face_up):
    v, idx = _floor(face_up)
    alb = np.array([[0.7, 0.5, 0.3]] * 2, np.float32)
    sc = orc.Scene(v, idx, albedo=alb)
    R = 1024
    to_light = np.array([0.3, 0.9, -0.2])
    to_light /= np.linalg.norm(to_light)
    Lvec = (-to_light).astype(np.float32)             # direction the light travels
    E = np.array([3.0, 2.5, 2.0])
    p = _params(_probes(), R, lights=[(vlb.LIGHT_DIRECTIONAL, Lvec.tolist(), E.tolist())])
    r = sc.bake(p, mode64=True, want_radiance=True, want_hits=True)
    d = orc.fibonacci(R)
    down = d[:, 1] < 0
    t, P = _floor_hits(_probes(), d)
    inside = down[None, :] & (np.abs(P[..., 0]) < L_FLOOR - 1) & (np.abs(P[..., 2]) < L_FLOOR - 1)
    hits = r["hits"] >= 0
    assert (hits[inside]).all() and not hits[:, ~down].any() and inside.sum() > 0.9 * 3 * down.sum()
    cos = float(-Lvec.astype(np.float64)[1])           # N = +y facing the probes above
    want = alb[0].astype(np.float64) * E * cos / np.pi
    rad = r["radiance"].astype(np.float64)
    np.testing.assert_allclose(rad[hits], np.broadcast_to(want, rad[hits].shape), rtol=1e-6)
    assert (rad[~hits] == 0).all()                     # black sky
    np.testing.assert_allclose(r["coeffs"], _project(r["radiance"], d), rtol=2e-7, atol=1e-9)


def test_lambert_point_light_inverse_square_and_shadow():
    v, idx = _floor(False)
    # occluder: a 1 x 1 square at y = 2.5 under the light, facing any way
    sq = np.array([[-0.5, 2.5, -0.5], [0.5, 2.5, -0.5], [0.5, 2.5, 0.5], [-0.5, 2.5, 0.5]], np.float32)
    xyz = np.concatenate([v, sq])
    idx = np.concatenate([idx, np.array([[4, 5, 6], [4, 6, 7]], np.uint32)])
    alb = np.array([[0.6, 0.6, 0.6]] * 4, np.float32)
    sc = orc.Scene(xyz, idx, albedo=alb)
    R = 2048
    Lpos = np.array([0.1, 4.0, 0.2], np.float32)
    I = np.array([40.0, 30.0, 20.0])
    probes = np.array([[2.0, 1.0, 0.0], [-1.5, 1.2, 1.0]], np.float32)
    p = _params(probes, R, lights=[(vlb.LIGHT_POINT, Lpos.tolist(), I.tolist())])
    r = sc.bake(p, mode64=True, want_radiance=True, want_hits=True)
    d = orc.fibonacci(R)
    t, P = _floor_hits(probes, d)
    Lp = Lpos.astype(np.float64)
    floor_hit = r["hits"] <= 1
    assert floor_hit.sum() > 0.3 * floor_hit.size
    rad = r["radiance"].astype(np.float64)
    tested = 0
    for pi in range(probes.shape[0]):
        for ri in np.nonzero((r["hits"][pi] >= 0) & (r["hits"][pi] <= 1))[0]:
            Ph = P[pi, ri]
            Lv = Lp - Ph
            dist = np.linalg.norm(Lv)
            cos = Lv[1] / dist
            # shadow: the segment from the floor point to the light crosses y = 2.5
            s = (2.5 - Ph[1]) / Lv[1]
            q = Ph + s * Lv
            if abs(abs(q[0]) - 0.5) < 1e-4 or abs(abs(q[2]) - 0.5) < 1e-4:
                continue                                # grazing the square's edge
            shadow = abs(q[0]) < 0.5 and abs(q[2]) < 0.5
            want = 0.0 if shadow else float(np.float32(0.6)) * I * cos / (np.pi * dist * dist)
            # the float32 hit point (~1e-5 m here) enters through 1/d^2
            np.testing.assert_allclose(rad[pi, ri], np.broadcast_to(want, (3,)), rtol=5e-5, atol=1e-9)
            tested += 1
    assert tested > 500
    # the square's underside faces the probes and turns away from the light
    sqh = r["hits"] >= 2
    assert sqh.any() and (rad[sqh] == 0).all()


def test_reference_shading_on_floor():
    v, idx = _floor(True)                              # geometric normal +y
    alb = np.array([[0.8, 0.4, 0.2]] * 2, np.float32)
    sc = orc.Scene(v, idx, albedo=alb)
    R = 1024
    Lpos = np.array([0.5, 3.0, -0.3], np.float32)
    amb, cd, cs, gl = 0.1, 0.6, 0.4, 16.0
    probes = _probes()
    p = _params(probes, R, shade_mode=vlb.SHADE_REFERENCE, ref_light=Lpos.tolist(), ref_bias=0.005,
                ref_ambient=amb, ref_cd=cd, ref_cs=cs, ref_gloss=gl, ref_flags=0)
    r = sc.bake(p, mode64=True, want_radiance=True, want_hits=True)
    d = orc.fibonacci(R)
    t, P = _floor_hits(probes, d)
    hit = r["hits"] >= 0
    assert not hit[:, d[:, 1] >= 0].any() and hit.sum() > 0.45 * hit.size
    a = alb[0].astype(np.float64)
    N = np.array([0.0, 1.0, 0.0])
    for pi in range(probes.shape[0]):
        h = hit[pi]
        Ph = P[pi, h]
        Lv = Lpos.astype(np.float64) - Ph
        Ln = Lv / np.linalg.norm(Lv, axis=1, keepdims=True)
        sdn = np.maximum(Ln @ N, 0.0)
        Rf = Ln - 2.0 * (Ln @ N)[:, None] * N           # reflect(L, N)
        rd = np.maximum((Rf * d[h].astype(np.float64)).sum(axis=1), 0.0)
        want = a[None, :] * (amb + cd * sdn + cs * rd ** gl)[:, None]
        # the float32 hit point enters through (R.d)^16
        np.testing.assert_allclose(r["radiance"][pi, h].astype(np.float64), want, rtol=5e-5, atol=1e-9)
    # misses see the configuration's sky, zero here
    assert (r["radiance"][~hit] == 0).all()
