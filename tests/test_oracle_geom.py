"""Pins of the oracle's ray queries (CPU): analytic hits, watertightness,
closed surfaces, BVH == brute force, shadow query == closest-hit distance."""
import numpy as np
import pytest

import orc
import vlb


def _unit(n, seed):
    d = np.random.default_rng(seed).normal(size=(n, 3))
    return (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)


def test_plane_hit_analytic():
    # floor quad y = 2 over [0,10]^2 as two triangles
    xyz = np.array([[0, 2, 0], [10, 2, 0], [10, 2, 10], [0, 2, 10]], np.float32)
    idx = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    s = orc.Scene(xyz, idx, bvh=False)
    org = np.array([[3, 7, 4], [1, 3, 9], [5, -1, 5]], np.float32)
    d = np.array([[0, -1, 0], [0, -1, 0], [0, 1, 0]], np.float32)
    p, t, u, v = s.trace_closest(org, d, 0.0, 100.0)
    assert (p >= 0).all()
    np.testing.assert_array_equal(t, np.array([5, 1, 3], np.float32))
    # barycentrics reproduce the hit point: P = (1-u-v) v0 + u v1 + v2 v
    for i in range(3):
        tri = xyz[idx[p[i]]].astype(np.float64)
        P = (1 - u[i] - v[i]) * tri[0] + u[i] * tri[1] + v[i] * tri[2]
        np.testing.assert_allclose(P, org[i] + d[i] * t[i], atol=1e-5)


def test_interval_is_inclusive_and_respects_tmin_tmax():
    xyz = np.array([[-1, -1, 4], [1, -1, 4], [0, 1, 4]], np.float32)
    s = orc.Scene(xyz, np.array([[0, 1, 2]], np.uint32), bvh=False)
    o, d = np.zeros((1, 3), np.float32), np.array([[0, 0, 1]], np.float32)
    assert s.trace_closest(o, d, 0.0, 4.0)[0][0] == 0        # t == tmax accepted
    assert s.trace_closest(o, d, 4.0, 10.0)[0][0] == 0       # t == tmin accepted
    assert s.trace_closest(o, d, 0.0, 3.999)[0][0] == -1
    assert s.trace_closest(o, d, 4.001, 10.0)[0][0] == -1
    assert s.trace_occluded(o, d, 4.0)[0] == 1 and s.trace_occluded(o, d, 3.5)[0] == 0


def test_shared_edge_and_vertex_rays_never_leak():
    # a fan of 8 triangles around the origin in the plane z = 0: rays aimed
    # exactly at the shared edges and the shared centre vertex must hit
    ang = np.linspace(0, 2 * np.pi, 9)[:-1]
    ring = np.stack([np.cos(ang), np.sin(ang), np.zeros(8)], 1)
    xyz = np.vstack([[0, 0, 0], ring]).astype(np.float32)
    idx = np.array([[0, 1 + i, 1 + (i + 1) % 8] for i in range(8)], np.uint32)
    s = orc.Scene(xyz, idx)
    targets = [np.zeros(3)] + [ring[i] * f for i in range(8) for f in (0.25, 0.5, 0.75)]
    targets = np.array(targets, np.float64)
    rng = np.random.default_rng(0)
    org = (targets + rng.uniform(-3, 3, size=targets.shape) * [1, 1, 0] + [0, 0, 5]).astype(np.float32)
    d = (targets - org).astype(np.float32)
    p = s.trace_closest(org, d, 0.0, 10.0)[0]
    assert (p >= 0).all()
    # a hit on a shared edge goes to the lowest prim id among the hit ones
    pb = s.trace_closest(org, d, 0.0, 10.0, brute=True)[0]
    assert np.array_equal(p, pb)


def test_rays_from_inside_closed_icosphere_never_escape():
    sc = vlb.generate_scene(vlb.SCENE_SPHERES, 3, 1, 4)  # one sphere, subdiv 4
    s = orc.Scene.from_dict(sc)
    c = sc["xyz"].astype(np.float64).mean(axis=0)
    d = _unit(20000, 1)
    org = np.repeat(c[None].astype(np.float32), len(d), 0)
    p, t, _, _ = s.trace_closest(org, d, 0.0, 1e4)
    assert (p >= 0).all()
    r = np.linalg.norm(sc["xyz"] - c, axis=1)
    assert t.min() > 0.95 * r.min() and t.max() <= r.max() * 1.0001


def test_box_axis_distances():
    sc = vlb.generate_scene(vlb.SCENE_CUBE, 0)
    s = orc.Scene.from_dict(sc)
    o = np.zeros((6, 3), np.float32)
    d = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], np.float32)
    p, t, _, _ = s.trace_closest(o, d, 0.0, 10.0)
    assert (p >= 0).all() and np.array_equal(t, np.ones(6, np.float32))


def test_bvh_equals_brute_force():
    sc = vlb.generate_scene(vlb.SCENE_ROOM, 5)
    s = orc.Scene.from_dict(sc)
    rng = np.random.default_rng(2)
    org = rng.uniform([0.1, 0.1, 0.1], [9.9, 3.9, 9.9], size=(4000, 3)).astype(np.float32)
    d = _unit(4000, 3)
    a = s.trace_closest(org, d, 0.001, 1e4)
    b = s.trace_closest(org, d, 0.001, 1e4, brute=True)
    for x, y in zip(a, b):
        assert np.array_equal(np.asarray(x).view(np.uint32), np.asarray(y).view(np.uint32))
    tmax = rng.uniform(0.01, 15, size=4000).astype(np.float32)
    assert np.array_equal(s.trace_occluded(org, d, tmax), s.trace_occluded(org, d, tmax, brute=True))


def test_occluded_agrees_with_closest_distance():
    sc = vlb.generate_scene(vlb.SCENE_ROOM, 6)
    s = orc.Scene.from_dict(sc)
    rng = np.random.default_rng(4)
    org = rng.uniform([0.1, 0.1, 0.1], [9.9, 3.9, 9.9], size=(3000, 3)).astype(np.float32)
    d = _unit(3000, 5)
    p, t, _, _ = s.trace_closest(org, d, 0.0, 1e4)
    tmax = rng.uniform(0.01, 15, size=3000).astype(np.float32)
    occ = s.trace_occluded(org, d, tmax, 0.0)
    expect = (p >= 0) & (t <= tmax)
    assert np.array_equal(occ.astype(bool), expect)


def test_scene_generators_are_deterministic_and_sized():
    a = vlb.generate_scene(vlb.SCENE_ROOM, 2)
    b = vlb.generate_scene(vlb.SCENE_ROOM, 2)
    c = vlb.generate_scene(vlb.SCENE_ROOM, 3)
    assert a["idx"].shape[0] == 11018
    for k in ("xyz", "idx", "normals", "albedo"):
        assert np.array_equal(a[k], b[k])
    assert not np.array_equal(a["xyz"], c["xyz"])
    t = vlb.generate_scene(vlb.SCENE_TERRAIN, 3, 64, 2)
    assert t["idx"].shape[0] == 2 * 63 * 63 + 2 * 20480
    assert vlb.generate_scene(vlb.SCENE_CUBE, 0)["idx"].shape[0] == 12
    for s in (a, t):
        assert (s["idx"] < s["xyz"].shape[0]).all()
        n = np.linalg.norm(s["normals"], axis=1)
        assert np.abs(n - 1).max() < 1e-5
        assert (s["albedo"] >= 0).all() and (s["albedo"] <= 1).all()


@pytest.mark.slow
def test_full_size_scenes_triangle_counts():
    assert vlb.generate_scene(vlb.SCENE_TERRAIN, 3)["idx"].shape[0] == 1154050
    city = vlb.generate_scene(vlb.SCENE_CITY, 5)["idx"].shape[0]
    assert 1.8e6 < city < 2.3e6
