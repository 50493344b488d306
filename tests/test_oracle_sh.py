"""Pins of the oracle's SH basis, direction sets and projections (CPU).

Each check compares the oracle with something independent of its formulas:
scipy's complex spherical harmonics, quadrature identities, Funk-Hecke closed
forms of zonal fields, or exact symmetry -- so a dropped term, a wrong sign or
a transposed index in the oracle fails here.
"""
import numpy as np
import pytest
from scipy import integrate, special

import orc
import vlb


def _real_sh_scipy(l, m, d):
    """Real SH of sh_common.h's convention built from scipy's complex Y_l^m
    (which carries the Condon-Shortley phase): m>0 -> sqrt2 Re Y_l^m,
    m<0 -> sqrt2 Im Y_l^|m|, m=0 -> Y_l^0."""
    x, y, z = d[:, 0], d[:, 1], d[:, 2]
    theta = np.arccos(np.clip(z, -1, 1))
    phi = np.arctan2(y, x)
    if hasattr(special, "sph_harm_y"):
        Y = special.sph_harm_y(l, abs(m), theta, phi)
    else:  # older scipy: sph_harm(m, l, azimuth, polar)
        Y = special.sph_harm(abs(m), l, phi, theta)
    if m > 0:
        return np.sqrt(2.0) * Y.real
    if m < 0:
        return np.sqrt(2.0) * Y.imag
    return Y.real


def _unit(n, seed):
    d = np.random.default_rng(seed).normal(size=(n, 3))
    return d / np.linalg.norm(d, axis=1, keepdims=True)


def test_exact_basis_matches_scipy():
    d = _unit(200, 1)
    got = orc.sh_eval(d, K=25, exact=True)
    for l in range(5):
        for m in range(-l, l + 1):
            ref = _real_sh_scipy(l, m, d)
            np.testing.assert_allclose(got[:, l * (l + 1) + m], ref, rtol=1e-11, atol=1e-12,
                                       err_msg=f"l={l} m={m}")


def test_literal_constants_are_rounded_exact_constants():
    # sh_common.h's literals are the exact normalisations rounded to 6
    # significant figures: lit/exact is the same constant for every direction
    # (same polynomial), within 1.4e-6 of 1 for l <= 3 and 5e-6 for l = 4
    d = _unit(500, 2)
    lit = orc.sh_eval(d, K=25, exact=False)
    ex = orc.sh_eval(d, K=25, exact=True)
    ok = np.abs(ex) > 1e-3
    for k in range(25):
        r = lit[ok[:, k], k] / ex[ok[:, k], k]
        assert r.max() - r.min() < 1e-12, k
        bound = 1.4e-6 if k < 16 else 5e-6
        assert abs(r.mean() - 1.0) <= bound, (k, r.mean() - 1.0)


def test_orthonormality_product_quadrature():
    # Gauss-Legendre in cos(theta) x uniform phi is exact for these degrees
    n = 16
    t, wt = np.polynomial.legendre.leggauss(n)
    phi = (np.arange(2 * n) + 0.5) * (2 * np.pi / (2 * n))
    T, PH = np.meshgrid(t, phi, indexing="ij")
    W = np.outer(wt, np.full(2 * n, 2 * np.pi / (2 * n))).ravel()
    s = np.sqrt(1 - T ** 2)
    d = np.stack([s * np.cos(PH), s * np.sin(PH), T], -1).reshape(-1, 3)
    Y = orc.sh_eval(d, K=25, exact=True)
    G = (Y * W[:, None]).T @ Y
    np.testing.assert_allclose(G, np.eye(25), atol=1e-12)


def test_fibonacci_set():
    for R in (1, 7, 1024, 4096):
        d = orc.fibonacci(R).astype(np.float64)
        assert np.abs(np.linalg.norm(d, axis=1) - 1).max() < 1e-6
        assert np.array_equal(orc.fibonacci(R), vlb.directions(vlb.DIRS_FIBONACCI, R)[0])
    d = orc.fibonacci(4096).astype(np.float64)
    Y = orc.sh_eval(d[:, [0, 2, 1]], K=9)
    G = (Y.T @ Y) * (4 * np.pi / 4096)
    np.testing.assert_allclose(G, np.eye(9), atol=2e-3)


def test_equirect_dirs_and_weights():
    xyz, w = orc.equirect_dirs(314, 100)
    xyz2, w2 = vlb.directions(vlb.DIRS_EQUIRECT, 0, 314, 100)
    assert np.array_equal(xyz, xyz2) and np.array_equal(w, w2)
    assert abs(w.astype(np.float64).sum() - 4 * np.pi) < 2e-3
    # env_map.rgen: the first row looks straight up (+y world) with .xzy
    assert xyz[0, 1] > 0.999
    assert abs(np.linalg.norm(xyz.astype(np.float64), axis=1) - 1).max() < 1e-6


def _funk_hecke(f, lmax=3):
    """lambda_l = 2 pi int_{-1}^{1} f(t) P_l(t) dt."""
    out = []
    for l in range(lmax + 1):
        pts = [0.0] if l >= 0 else None
        v, _ = integrate.quad(lambda t: f(t) * special.eval_legendre(l, t), -1, 1, points=pts,
                              limit=400, epsabs=1e-13, epsrel=1e-12)
        out.append(2 * np.pi * v)
    return np.array(out)


def _zonal_coeffs(lam, s_sh, K):
    Y = orc.sh_eval(np.asarray([s_sh]), K=K, exact=True)[0]
    ls = np.array([int(np.floor(np.sqrt(k))) for k in range(K)])
    return lam[ls] * Y


def test_skybox_constant_sky():
    cube = np.zeros((6, 32, 32, 4), np.float32)
    cube[..., 0], cube[..., 1], cube[..., 2] = 1.0, 0.5, 2.0
    c = orc.skybox_sh(cube, K=16)
    np.testing.assert_allclose(c[0], 2 * np.sqrt(np.pi) * np.array([1.0, 0.5, 2.0]), rtol=1e-12)
    assert np.abs(c[1:]).max() < 1e-12


def test_skybox_sun_lobe_closed_form():
    # zonal lobe 50 exp(200 (d.s - 1)) projects to lambda_l Y_lm(s)
    s = np.array([0.3, 0.8, -0.52])
    s /= np.linalg.norm(s)
    cube = orc.procedural_cubemap(256, s, zenith=0.0, horizon=0.0, sun_rgb=50.0, k=200.0)
    got = orc.skybox_sh(cube, K=16)[:, 0]
    lam = _funk_hecke(lambda t: 50.0 * np.exp(200.0 * (t - 1.0)))
    ref = _zonal_coeffs(lam, s[[0, 2, 1]], 16)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 2e-3


def test_skybox_gradient_closed_form():
    # sky gradient lerp(0.5, 2.0, max(d.y, 0)): zonal about world up = SH +z
    cube = orc.procedural_cubemap(128, [0, 1, 0], zenith=2.0, horizon=0.5, sun_rgb=0.0)
    got = orc.skybox_sh(cube, K=16)[:, 0]
    lam = _funk_hecke(lambda t: 0.5 + 1.5 * max(t, 0.0))
    ref = _zonal_coeffs(lam, np.array([0.0, 0.0, 1.0]), 16)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-4


@pytest.mark.parametrize("variant", [vlb.EQUIRECT_SH_COMP, vlb.EQUIRECT_SKYBOX_COMP])
def test_equirect_projection_constant_and_linear(variant):
    W, H = 628, 200
    img = np.full((H, W, 4), 255, np.uint8)
    c = orc.equirect_project(img, variant)
    # sh.comp's constant c00 = 0.282095 * sum(w) with the midpoint rule
    assert abs(c[0, 0] - 0.282095 * 4 * np.pi) < 1e-4 * 0.282095 * 4 * np.pi
    assert np.abs(c[1:]).max() < 1e-3
    # an image whose red channel is (1 + cos theta)/2 (theta from the image row):
    # in both variants theta is the SH polar angle (sh.comp SH(dir) has
    # z = cos theta; skybox_sh.comp's dir.xzy puts cos theta in SH y)
    theta = np.pi * (np.arange(H) + 0.5) / H
    val = np.rint(255 * (1 + np.cos(theta)) / 2).astype(np.uint8)
    img = np.zeros((H, W, 4), np.uint8)
    img[..., 0] = val[:, None]
    c = orc.equirect_project(img, variant)
    k_axis = 2 if variant == vlb.EQUIRECT_SH_COMP else 1  # Y10 (z) vs Y1-1 (-y)
    sign = 1.0 if variant == vlb.EQUIRECT_SH_COMP else -1.0
    ref = sign * 0.488603 * 0.5 * 4 * np.pi / 3
    assert abs(c[k_axis, 0] - ref) < 1e-2 * abs(ref)
    assert abs(c[0, 0] - 0.282095 * 0.5 * 4 * np.pi) < 1e-2 * 0.282095 * 2 * np.pi


def test_reconstruct_projection_round_trip():
    # sh_sum.comp reconstruction of a band-limited field, projected back with
    # sh.comp, returns the coefficients (up to RGBA8 quantisation)
    rng = np.random.default_rng(3)
    c = np.zeros((16, 3), np.float32)
    c[0] = 1.5e-3                      # 1250 * field stays inside (0, 1):
    c[1:] = rng.uniform(-3e-5, 3e-5, size=(15, 3))  # no clamping on store
    img = orc.sh_reconstruct(c, 628, 200)
    assert 0 < img[..., :3].min() and img[..., :3].max() < 255
    back = orc.equirect_project(img, vlb.EQUIRECT_SH_COMP) / 1250.0
    err = np.linalg.norm(back - c, axis=0) / np.linalg.norm(c, axis=0)
    assert err.max() < 5e-3


def test_srgb_and_unorm8_helpers():
    assert orc.L.orc_srgb(0.001) == pytest.approx(0.001 * 12.92, rel=1e-6)
    assert orc.L.orc_srgb(1.0) == pytest.approx(1.0, rel=1e-6)
    assert orc.L.orc_srgb(0.5) == pytest.approx(1.055 * 0.5 ** (1 / 2.4) - 0.055, rel=1e-6)
    for x, q in [(0.0, 0), (1.0, 255), (0.5, 128), (-3.0, 0), (7.0, 255), (0.5 / 255, 0), (1.5 / 255, 2)]:
        assert orc.L.orc_unorm8(x) == pytest.approx(q / 255.0, abs=1e-7)
