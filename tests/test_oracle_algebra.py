"""Pins of the oracle's plane algebra (pmhstereo.cl:50-153) and image
pre-processing (pmhstereo.cl:737-831) against closed forms and FP64 stencils."""
import numpy as np
import pytest


def plane_from_nd(n, z, x, y):
    """Left-view plane with normal n at depth z through pixel (x, y), FP64."""
    nx, ny, nz = n
    return np.array([1 + nx / nz, ny / nz, -((nx * x + ny * y) / nz + z)])


@pytest.mark.parametrize("u", [(0.3, 0.5, 0.5), (0.9, 0.1, 0.8), (0.0, 1.0, 0.0), (0.5, 0.2, 0.93)])
def test_init_closed_form(orc, u):
    x, y, dmax = 37.0, 11.0, 64.0
    s = orc.state_init(u, x, y, 1, 1.0, dmax)
    z = u[0] * dmax
    nx = np.clip((u[1] - 0.5) * 0.95, -0.95, 0.95)
    ny = np.clip((u[2] - 0.5) * 0.95, -0.95, 0.95)
    f = max(np.hypot(nx, ny) + 0.001, 1.0)
    nx, ny = nx / f, ny / f
    nz = np.sqrt(1 - nx * nx - ny * ny)
    np.testing.assert_allclose(s[:3], plane_from_nd((nx, ny, nz), z, x, y), rtol=2e-6, atol=2e-5)
    # the plane maps its own pixel to x - z
    assert abs(orc.state_transform(s, x, y) - (x - z)) < 1e-4
    # and ref_normal_depth recovers (n, z)
    nd = orc.state_to_ref_normal_depth(s, x, y, 1)
    np.testing.assert_allclose(nd, [nx, ny, nz, z], rtol=1e-5, atol=1e-4)


def test_viewprop_is_involution_and_inverts_mapping(orc):
    rng = np.random.default_rng(1)
    for _ in range(100):
        s = np.array([rng.uniform(0.5, 1.5), rng.uniform(-0.3, 0.3), rng.uniform(-80, 10), 0],
                     np.float32)
        v = orc.state_viewprop(s)
        np.testing.assert_allclose(orc.state_viewprop(v)[:3], s[:3], rtol=1e-6, atol=1e-5)
        x, y = rng.uniform(0, 300), rng.uniform(0, 200)
        xp = orc.state_transform(s, x, y)
        assert abs(orc.state_transform(v, xp, y) - x) < 1e-3
        assert v[3] == 0.0


def test_right_view_init_is_viewprop_of_left(orc):
    u = (0.4, 0.3, 0.6)
    l = orc.state_init(u, 20.0, 5.0, 1, 2.0, 100.0)
    r = orc.state_init(u, 20.0, 5.0, 0, 2.0, 100.0)
    np.testing.assert_array_equal(r, orc.state_viewprop(l))


def test_refine_zero_perturbation_keeps_plane(orc):
    s = orc.state_init((0.3, 0.45, 0.6), 50.0, 20.0, 1, 1.0, 64.0)
    r = orc.state_refine((0.5, 0.5, 0.5), s, 50.0, 20.0, 64.0, 1)
    np.testing.assert_allclose(r[:3], s[:3], rtol=1e-5, atol=1e-3)


def test_refine_perturbs_depth_within_one(orc):
    s = orc.state_init((0.3, 0.5, 0.5), 50.0, 20.0, 1, 1.0, 64.0)
    z0 = 50.0 - orc.state_transform(s, 50.0, 20.0)
    r = orc.state_refine((1.0, 0.5, 0.5), s, 50.0, 20.0, 64.0, 1)
    z1 = 50.0 - orc.state_transform(r, 50.0, 20.0)
    assert abs((z1 - z0) - 1.0) < 1e-3


def test_refine_right_view_quirk(orc):
    # pmhstereo.cl:114-115: the right view returns viewprop((1,0,0,0) - viewprop(s))
    s = np.array([1.02, 0.05, 12.0, 0.3], np.float32)
    r = orc.state_refine((0.9, 0.1, 0.7), s, 10.0, 10.0, 64.0, 0)
    v = orc.state_viewprop(s)
    t = np.array([1 - v[0], -v[1], -v[2], 0], np.float32)
    np.testing.assert_array_equal(r, orc.state_viewprop(t))


def test_to_normal_fronto_parallel(orc):
    n = orc.state_to_normal(np.array([1.0, 0.0, -5.0, 0.0], np.float32))
    np.testing.assert_allclose(n, [0, 0, 1], atol=1e-7)


def test_nonfinite_ref_normal_depth(orc):
    nd = orc.state_to_ref_normal_depth(np.array([np.inf, 0, 0, 0], np.float32), 1.0, 1.0, 1)
    np.testing.assert_array_equal(nd, [0, 0, 1, 0])


def _gray64(img):
    I = img.astype(np.float64)
    return I  # 0.2126 + 0.7152 + 0.0722 = 1 up to float rounding


def test_gradxy_fp64_stencil(orc):
    rng = np.random.default_rng(3)
    img = rng.random((23, 31)).astype(np.float32)
    g = orc.gradxy(img)
    P = np.pad(_gray64(img), 1)
    c = P[1:-1, 1:-1]
    xm, xp = P[1:-1, :-2], P[1:-1, 2:]
    ym, yp = P[:-2, 1:-1], P[2:, 1:-1]
    mm, pm, mp, pp = P[:-2, :-2], P[:-2, 2:], P[2:, :-2], P[2:, 2:]
    dx = 0.5 * (xp - xm) + 0.25 * (pm - mm) + 0.25 * (pp - mp)
    dy = 0.5 * (yp - ym) + 0.25 * (mp - mm) + 0.25 * (pp - pm)
    ref = np.stack([dx, dy, pp - mm, mp - pm], -1)
    np.testing.assert_allclose(g, ref, atol=2e-6)
    assert c.shape == g.shape[:2]


def test_weight_fp64_with_zero_dy(orc):
    rng = np.random.default_rng(4)
    img = rng.random((17, 29)).astype(np.float32)
    w = orc.weight(img)
    P = np.pad(_gray64(img), 1)
    dx = P[1:-1, 2:] - P[1:-1, :-2]
    ref = np.exp(-3 * np.abs(dx) ** 0.8) + 1e-4
    np.testing.assert_allclose(w, ref, rtol=2e-5, atol=1e-6)


def test_factor_table(orc):
    F = orc.factor_table(17)
    dy, dx = np.mgrid[-17:18, -17:18]
    t = np.clip(np.hypot(dx, dy) / 26.0, 0, 1)
    np.testing.assert_allclose(F, t * t * (3 - 2 * t) * 39 + 5, rtol=1e-6)
    assert F[17, 17] == 5.0
