"""Pins of the oracle's probe-grid lookup (CPU).

The checks compare orc_probe_lookup with facts fixed independently of its
code: weights that sum to one, exact trilinear reproduction of a linear
field, the clamped-cosine band factors A_l / pi = 2 int_0^1 t P_l(t) dt
integrated numerically by scipy (with the bake's SH frame: a wrong swizzle or
factor changes the answer), the breadth-first probe order of orc_probe_grid,
and a wall that hides half the probes.
"""
import numpy as np
import pytest
from scipy import integrate, special

import orc
import vlb


def _unit(n, seed):
    d = np.random.default_rng(seed).normal(size=(n, 3))
    return (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)


def _grid(count=(3, 2, 4), step=(1.5, 2.0, 0.75), origin=(0.5, -1.0, 2.0), K=9, order=None):
    return vlb.grid_desc(origin, step, count, order=order, num_coefs=K)


def _inside(g, n, seed):
    rng = np.random.default_rng(seed)
    o = np.array(g.origin[:], np.float64)
    ext = np.array(g.step[:]) * (np.array(g.count[:]) - 1)
    return (o + rng.uniform(0.0, 1.0, (n, 3)) * ext).astype(np.float32)


def _band_factor(l):
    """A_l / pi of the clamped cosine: 2 int_0^1 t P_l(t) dt (Ramamoorthi-Hanrahan)."""
    v, _ = integrate.quad(lambda t: t * special.eval_legendre(l, t), 0.0, 1.0, epsabs=1e-14)
    return 2.0 * v


@pytest.mark.parametrize("mode", [vlb.LOOKUP_REFERENCE, vlb.LOOKUP_IRRADIANCE])
def test_equal_probes_give_that_probe(mode):
    g = _grid()
    P = int(np.prod(g.count[:]))
    c1 = np.random.default_rng(1).normal(size=(9, 3))
    coeffs = np.broadcast_to(c1, (P, 9, 3)).astype(np.float32)
    pts, nrm = _inside(g, 50, 2), _unit(50, 3)
    got = orc.probe_lookup(None, g, coeffs, vlb.lookup_params(mode, False), pts, nrm)
    Y = orc.sh_eval(nrm.astype(np.float64)[:, [0, 2, 1]], K=9, exact=True)
    band = np.ones(9) if mode == vlb.LOOKUP_REFERENCE else np.array([1.0] + [2 / 3] * 3 + [0.25] * 5)
    ref = (Y * band) @ c1.astype(np.float32).astype(np.float64)
    np.testing.assert_allclose(got, ref, rtol=1e-6, atol=1e-7)


def test_trilinear_reproduces_a_linear_field():
    g = _grid()
    pos = orc.probe_grid(g.origin[:], [g.origin[a] + g.step[a] * (g.count[a] - 1) for a in range(3)],
                         list(g.count[:]), vlb.ORDER_LINEAR)[0].astype(np.float64)
    a, b = np.array([0.3, -1.0, 2.0]), np.array([[0.5, -0.25, 1.0], [2.0, 0.0, -1.0], [0.1, 0.2, 0.3]])
    field = a + pos @ b                          # (P, 3) per channel
    coeffs = np.zeros((pos.shape[0], 9, 3), np.float32)
    y00 = 0.5 / np.sqrt(np.pi)
    coeffs[:, 0, :] = field / y00
    pts = _inside(g, 80, 4)
    got = orc.probe_lookup(None, g, coeffs, vlb.lookup_params(vlb.LOOKUP_IRRADIANCE, False), pts,
                           _unit(80, 5))
    np.testing.assert_allclose(got, a + pts.astype(np.float64) @ b, rtol=2e-5, atol=2e-5)


@pytest.mark.parametrize("K", [9, 16])
def test_band_factors_and_frame_closed_forms(K):
    g = _grid(count=(2, 2, 2), K=K)
    P = 8
    nrm = _unit(64, 6)
    pts = _inside(g, 64, 7)
    for l in range(4 if K == 16 else 3):
        for m in range(-l, l + 1):
            k = l * (l + 1) + m
            coeffs = np.zeros((P, K, 3), np.float32)
            coeffs[:, k, :] = 1.0
            got = orc.probe_lookup(None, g, coeffs, vlb.lookup_params(vlb.LOOKUP_IRRADIANCE, False),
                                   pts, nrm)[:, 0]
            # radiance Y_lm(d) convolves to A_l/pi * Y_lm(n), SH at world (x, z, y)
            Y = orc.sh_eval(nrm.astype(np.float64)[:, [0, 2, 1]], K=K, exact=True)[:, k]
            np.testing.assert_allclose(got, _band_factor(l) * Y, rtol=1e-6, atol=1e-7,
                                       err_msg=f"l={l} m={m}")
    # the l = 1 field L(d) = d.s (world s): irradiance / pi = (2/3) n.s
    s = np.array([0.3, 0.8, -0.52])
    s /= np.linalg.norm(s)
    c = np.zeros((P, K, 3), np.float32)
    ssh = s[[0, 2, 1]]                    # SH frame (x, z, y)
    c[:, 2, 0] = ssh[2] / 0.48860251190291992
    c[:, 3, 0] = -ssh[0] / 0.48860251190291992
    c[:, 1, 0] = -ssh[1] / 0.48860251190291992
    got = orc.probe_lookup(None, g, c, vlb.lookup_params(vlb.LOOKUP_IRRADIANCE, False), pts, nrm)[:, 0]
    np.testing.assert_allclose(got, (2.0 / 3.0) * (nrm.astype(np.float64) @ s), rtol=1e-6, atol=1e-7)
    got = orc.probe_lookup(None, g, c, vlb.lookup_params(vlb.LOOKUP_REFERENCE, False), pts, nrm)[:, 0]
    np.testing.assert_allclose(got, nrm.astype(np.float64) @ s, rtol=1e-6, atol=1e-7)


def test_bfs_order_matches_the_grid_generator():
    cnt = (4, 3, 5)
    b0, step = np.array([1.0, 2.0, -3.0], np.float32), np.array([0.5, 1.25, 2.0], np.float32)
    b1 = b0 + step * (np.array(cnt) - 1)
    pos, st = orc.probe_grid(b0, b1, list(cnt), vlb.ORDER_REFERENCE_BFS)
    g = vlb.grid_desc(b0, st, cnt, order=vlb.ORDER_REFERENCE_BFS)
    # probe i carries its own (x, y, z) grid coordinates, recovered from its position
    ijk = np.rint((pos.astype(np.float64) - b0) / st).astype(int)
    coeffs = np.zeros((pos.shape[0], 9, 3), np.float32)
    coeffs[:, 0, :] = ijk * (2.0 * np.sqrt(np.pi))
    # a point just inside each probe's cell corner sees that probe with weight ~1
    pts = pos + np.float32(1e-4) * np.sign(st)
    got = orc.probe_lookup(None, g, coeffs, vlb.lookup_params(vlb.LOOKUP_IRRADIANCE, False), pts,
                           _unit(len(pts), 8))
    np.testing.assert_allclose(got, ijk, atol=2e-3)
    # the linear order too
    posl, _ = orc.probe_grid(b0, b1, list(cnt), vlb.ORDER_LINEAR)
    gl = vlb.grid_desc(b0, st, cnt, order=vlb.ORDER_LINEAR)
    ijkl = np.rint((posl.astype(np.float64) - b0) / st).astype(int)
    cl = np.zeros((posl.shape[0], 9, 3), np.float32)
    cl[:, 0, :] = ijkl * (2.0 * np.sqrt(np.pi))
    got = orc.probe_lookup(None, gl, cl, vlb.lookup_params(vlb.LOOKUP_IRRADIANCE, False),
                           posl + np.float32(1e-4) * np.sign(st), _unit(len(posl), 9))
    np.testing.assert_allclose(got, ijkl, atol=2e-3)


def test_reference_weights_and_visibility_wall():
    # 2x2x2 grid of step 1 at the origin, a wall at x = 0.5 covering the cell
    g = vlb.grid_desc((0, 0, 0), (1, 1, 1), (2, 2, 2))
    xyz = np.array([[0.5, -1, -1], [0.5, 2, -1], [0.5, 2, 2], [0.5, -1, 2]], np.float32)
    idx = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    wall = orc.Scene(xyz, idx)
    coeffs = np.zeros((8, 9, 3), np.float32)
    lin = np.arange(8)
    x = lin % 2
    coeffs[:, 0, 0] = np.where(x == 0, 1.0, 100.0) * 2.0 * np.sqrt(np.pi)
    p = np.array([[0.25, 0.4, 0.6]], np.float32)
    n = np.array([[-1.0, 0.0, 0.0]], np.float32)
    lp = vlb.lookup_params(vlb.LOOKUP_REFERENCE, True, 1e-3)
    got = orc.probe_lookup(wall, g, coeffs, lp, p, n)
    assert got[0, 0] == pytest.approx(1.0, rel=1e-6)   # only the x = 0 probes count
    # without visibility: weights |step| - distance over all 8 corners
    lp0 = vlb.lookup_params(vlb.LOOKUP_REFERENCE, False)
    got0 = orc.probe_lookup(wall, g, coeffs, lp0, p, n)[0, 0]
    corners = np.array([[(v >> 2) & 1, (v >> 1) & 1, v & 1] for v in range(8)], np.float64)
    w = np.maximum(np.sqrt(3.0) - np.linalg.norm(corners - p[0], axis=1), 0.0)
    val = np.where(corners[:, 0] == 0, 1.0, 100.0)
    assert got0 == pytest.approx((w * val).sum() / w.sum(), rel=1e-5)
