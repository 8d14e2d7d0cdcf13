"""Pins of the oracle's window cost (pmhstereo.cl:170-228) by closed forms.

Each case fixes the answer independently of the oracle's loop:
  - integer-disparity fronto-parallel plane on an exactly shifted pair -> 0;
  - constant offset between the views -> 0.15 * min(offset, 0.04);
  - linear ramp and a fractional disparity error -> 0.15 * slope * error
    (exercises the bilinear lookup);
  - row-wise offsets -> weighted mean with the support weights of the formula;
  - no valid tap / single tap / non-finite plane -> 1e5;
  - FP64 gradient against central finite differences.
"""
import numpy as np
import pytest

P = 5


def grads(orc, img):
    return orc.gradxy(img)


def test_exact_shift_gives_zero(orc):
    rng = np.random.default_rng(0)
    H, W, d = 30, 60, 7
    I1 = rng.random((H, W)).astype(np.float32)
    I2 = np.empty_like(I1)
    I2[:, : W - d] = I1[:, d:]
    I2[:, W - d:] = rng.random((H, d))
    G1, G2 = grads(orc, I1), grads(orc, I2)
    s = np.array([1.0, 0.0, -float(d), 0.0], np.float32)
    for (x, y) in [(20, 12), (30, 15), (25, 8)]:
        assert orc.cost(I1, G1, I2, G2, x, y, s, P) == 0.0


@pytest.mark.parametrize("off", [0.01, 0.025, 0.1])
def test_constant_offset(orc, off):
    H, W = 30, 40
    I1 = np.full((H, W), 0.4, np.float32)
    I2 = (I1 + np.float32(off)).astype(np.float32)
    G1, G2 = grads(orc, I1), grads(orc, I2)
    s = np.array([1.0, 0.0, 0.0, 0.0], np.float32)
    c = orc.cost(I1, G1, I2, G2, 20, 15, s, P)
    assert c == pytest.approx(0.15 * min(off, 0.04), rel=2e-5)


@pytest.mark.parametrize("delta", [0.3, 0.75, 1.4])
def test_ramp_fractional_disparity(orc, delta):
    H, W, slope, d = 30, 80, 0.01, 10
    x = np.arange(W, dtype=np.float64)
    I1 = np.tile(slope * x + 0.1, (H, 1)).astype(np.float32)
    I2 = np.tile(slope * (x + d) + 0.1, (H, 1)).astype(np.float32)
    G1, G2 = grads(orc, I1), grads(orc, I2)
    s = np.array([1.0, 0.0, -(d + delta), 0.0], np.float32)
    c = orc.cost(I1, G1, I2, G2, 40, 15, s, P)
    assert c == pytest.approx(0.15 * min(slope * delta, 0.04), rel=2e-4)


def test_support_weights_row_offsets(orc):
    rng = np.random.default_rng(5)
    H, W, y0 = 40, 50, 20
    I1 = (0.3 + 0.4 * rng.random((H, W))).astype(np.float32)
    off = np.where(np.arange(H) < y0, 0.01, 0.03)[:, None]
    I2 = (I1 + off).astype(np.float32)
    G1, G2 = grads(orc, I1), grads(orc, I2)
    s = np.array([1.0, 0.0, 0.0, 0.0], np.float32)
    x, y = 25, 18
    # x' = x exactly, so C is read off the arrays without interpolation: colour
    # 0.15*offset everywhere, plus the truncated jump of dy, dxy, dyx on rows
    # y0-1, y0 (taken from the gradient images, which are pinned separately)
    I1d, I2d = I1.astype(np.float64), I2.astype(np.float64)
    G1d, G2d = G1.astype(np.float64), G2.astype(np.float64)
    num = den = 0.0
    for qy in range(y - P, y + P + 1):
        for qx in range(x - P, x + P + 1):
            C = 0.15 * min(abs(I1d[qy, qx] - I2d[qy, qx]), 0.04) + \
                0.95 * np.minimum(np.abs(G1d[qy, qx] - G2d[qy, qx]), 0.01).sum()
            dx, dy = qx - x, qy - y
            t = min(np.hypot(dx, dy) / 26.0, 1.0)
            F = t * t * (3 - 2 * t) * 39 + 5
            w = np.exp(-3 * abs(I1d[y, x] - I1d[qy, qx]) * F)
            num += w * C
            den += w
    jump = np.minimum(np.abs(G1d[y0] - G2d[y0]), 0.01)
    assert np.allclose(jump[5:-5, 1:], 0.01)  # the rows the test is built on
    assert orc.cost(I1, G1, I2, G2, x, y, s, P) == pytest.approx(num / den, rel=1e-5)


def test_invalid_cases_give_1e5(orc):
    rng = np.random.default_rng(1)
    I1 = rng.random((20, 30)).astype(np.float32)
    I2 = rng.random((20, 30)).astype(np.float32)
    G1, G2 = grads(orc, I1), grads(orc, I2)
    far = np.array([1.0, 0.0, -1000.0, 0], np.float32)
    assert orc.cost(I1, G1, I2, G2, 10, 10, far, P) == 1e5
    nan = np.array([np.nan, 0.0, 0.0, 0], np.float32)
    assert orc.cost(I1, G1, I2, G2, 10, 10, nan, P) == 1e5
    inf = np.array([1.0, np.inf, 0.0, 0], np.float32)
    assert orc.cost(I1, G1, I2, G2, 10, 10, inf, P) == 1e5
    # a single tap carries weight exp(0) = 1 <= 1.1
    ok = np.array([1.0, 0.0, 0.0, 0], np.float32)
    assert orc.cost(I1, G1, I2, G2, 10, 10, ok, 0) == 1e5


def test_float_matches_fp64(orc):
    rng = np.random.default_rng(2)
    H, W = 40, 64
    I1 = rng.random((H, W)).astype(np.float32)
    I2 = rng.random((H, W)).astype(np.float32)
    G1, G2 = grads(orc, I1), grads(orc, I2)
    for _ in range(40):
        x, y = int(rng.integers(0, W)), int(rng.integers(0, H))
        s = np.array([rng.uniform(0.8, 1.2), rng.uniform(-0.1, 0.1), rng.uniform(-20, 5), 0],
                     np.float32)
        c32 = orc.cost(I1, G1, I2, G2, x, y, s, P)
        c64 = orc.cost_f64(I1, G1, I2, G2, x, y, s[:3].astype(np.float64), P)
        assert c32 == pytest.approx(c64, rel=1e-5)


def test_fp64_gradient_finite_differences(orc):
    rng = np.random.default_rng(7)
    H, W = 30, 50
    from scipy.ndimage import gaussian_filter
    I1 = gaussian_filter(rng.random((H, W)), 1.5).astype(np.float32)
    I2 = np.roll(I1, -4, axis=1) + 0.002 * rng.standard_normal((H, W)).astype(np.float32)
    I2 = I2.astype(np.float32)
    G1, G2 = grads(orc, I1), grads(orc, I2)
    checked = 0
    for _ in range(30):
        x, y = int(rng.integers(8, W - 8)), int(rng.integers(6, H - 6))
        s = np.array([1.0 + rng.uniform(-0.02, 0.02), rng.uniform(-0.02, 0.02),
                      -4.3 + rng.uniform(-0.5, 0.5)])
        c, dp, dI1, dG1, dI2, dG2 = orc.cost_grad_f64(I1, G1, I2, G2, x, y, s, P)
        for j, h in enumerate([1e-7, 1e-7, 1e-6]):
            e = np.zeros(3)
            e[j] = h
            fd1 = (orc.cost_f64(I1, G1, I2, G2, x, y, s + e, P)
                   - orc.cost_f64(I1, G1, I2, G2, x, y, s - e, P)) / (2 * h)
            fd2 = (orc.cost_f64(I1, G1, I2, G2, x, y, s + 10 * e, P)
                   - orc.cost_f64(I1, G1, I2, G2, x, y, s - 10 * e, P)) / (20 * h)
            if abs(fd1 - fd2) > 1e-3 * (abs(fd1) + 1e-6):
                continue  # a kink (truncation, cell edge) inside the stencil
            assert dp[j] == pytest.approx(fd1, rel=1e-4, abs=1e-7)
            checked += 1
        # image gradient at one texel of I2 inside the window footprint
        qy = y
        qx = int(s[0] * x + s[1] * y + s[2])
        h = 1e-4
        A = I2.copy(); A[qy, qx] += np.float32(h)
        B = I2.copy(); B[qy, qx] -= np.float32(h)
        hh = (float(A[qy, qx]) - float(B[qy, qx])) / 2
        fd = (orc.cost_f64(I1, G1, A, G2, x, y, s, P) - orc.cost_f64(I1, G1, B, G2, x, y, s, P)) / (2 * hh)
        fdb = (orc.cost_f64(I1, G1, I2, G2, x, y, s, P) - orc.cost_f64(I1, G1, B, G2, x, y, s, P)) / hh
        if abs(fd - fdb) < 1e-3 * (abs(fd) + 1e-6):
            assert dI2[qy, qx] == pytest.approx(fd, rel=1e-3, abs=1e-6)
            checked += 1
    assert checked > 40


def _rgba_pair(orc, rng, H, W, shift):
    """A textured RGBA pair (distinct channels, alpha random) and its
    gradient images; the right view is the left shifted by `shift` px."""
    from scipy.ndimage import gaussian_filter
    L = np.stack([gaussian_filter(rng.random((H, W)), 1.5) for _ in range(4)], -1)
    R = np.roll(L, -shift, axis=1) + 0.002 * rng.standard_normal((H, W, 4))
    L, R = L.astype(np.float32), R.astype(np.float32)
    return L, R, orc.gradxy(L), orc.gradxy(R)


def test_fp64_gradient_rgba_finite_differences(orc):
    """The RGBA FP64 gradient (colour sums over R, G, B) against central
    differences of the FP64 cost: the plane, single channels of I1 (a window
    texel and the centre pixel, which sets every support weight) and of I2."""
    rng = np.random.default_rng(17)
    H, W = 30, 50
    I1, I2, G1, G2 = _rgba_pair(orc, rng, H, W, 4)

    def fd_image(img, which, qy, qx, c, x, y, s, h=1e-4):
        A = img.copy(); A[qy, qx, c] += np.float32(h)
        B = img.copy(); B[qy, qx, c] -= np.float32(h)
        hh = (float(A[qy, qx, c]) - float(B[qy, qx, c])) / 2
        args = lambda im: (im, G1, I2, G2) if which == 1 else (I1, G1, im, G2)
        f = lambda im: orc.cost_f64(*args(im), x, y, s, P)
        fd = (f(A) - f(B)) / (2 * hh)
        fdb = (f(img) - f(B)) / hh
        return fd, abs(fd - fdb) < 1e-3 * (abs(fd) + 1e-6)

    checked = 0
    for _ in range(25):
        x, y = int(rng.integers(8, W - 8)), int(rng.integers(6, H - 6))
        s = np.array([1.0 + rng.uniform(-0.02, 0.02), rng.uniform(-0.02, 0.02),
                      -4.3 + rng.uniform(-0.5, 0.5)])
        c, dp, dI1, dG1, dI2, dG2 = orc.cost_grad_f64(I1, G1, I2, G2, x, y, s, P)
        assert dI1.shape == (H, W, 4) and not dI1[..., 3].any() and not dI2[..., 3].any()
        for j, h in enumerate([1e-7, 1e-7, 1e-6]):
            e = np.zeros(3)
            e[j] = h
            fd1 = (orc.cost_f64(I1, G1, I2, G2, x, y, s + e, P)
                   - orc.cost_f64(I1, G1, I2, G2, x, y, s - e, P)) / (2 * h)
            fd2 = (orc.cost_f64(I1, G1, I2, G2, x, y, s + 10 * e, P)
                   - orc.cost_f64(I1, G1, I2, G2, x, y, s - 10 * e, P)) / (20 * h)
            if abs(fd1 - fd2) > 1e-3 * (abs(fd1) + 1e-6):
                continue  # a kink (truncation, cell edge) inside the stencil
            assert dp[j] == pytest.approx(fd1, rel=1e-4, abs=1e-7)
            checked += 1
        ch = int(rng.integers(0, 3))
        qy, qx = y + int(rng.integers(-P, P + 1)), x + int(rng.integers(-P, P + 1))
        for (img, which, ty, tx, grad) in ((I1, 1, qy, qx, dI1), (I1, 1, y, x, dI1),
                                            (I2, 2, y, int(s[0] * x + s[1] * y + s[2]), dI2)):
            fd, smooth = fd_image(img, which, ty, tx, ch, x, y, s)
            if smooth:
                assert grad[ty, tx, ch] == pytest.approx(fd, rel=1e-3, abs=1e-6)
                checked += 1
    assert checked > 60


def test_fp64_rgba_with_equal_channels_is_gray(orc):
    """R = G = B = the gray image (alpha anything): the same cost and plane
    gradient, the gray image gradient split over the three colour channels
    (d/dgray = sum over R, G, B), none on alpha, the same gradient-image parts."""
    rng = np.random.default_rng(23)
    H, W = 24, 40
    from scipy.ndimage import gaussian_filter
    gL = gaussian_filter(rng.random((H, W)), 1.5).astype(np.float32)
    gR = (np.roll(gL, -3, axis=1) + 0.002 * rng.standard_normal((H, W))).astype(np.float32)
    cL = np.concatenate([np.repeat(gL[..., None], 3, -1), rng.random((H, W, 1))], -1)
    cR = np.concatenate([np.repeat(gR[..., None], 3, -1), rng.random((H, W, 1))], -1)
    cL, cR = cL.astype(np.float32), cR.astype(np.float32)
    G1, G2 = grads(orc, gL), grads(orc, gR)
    for _ in range(8):
        x, y = int(rng.integers(6, W - 6)), int(rng.integers(5, H - 5))
        s = np.array([1.0 + rng.uniform(-0.03, 0.03), rng.uniform(-0.03, 0.03),
                      -3.2 + rng.uniform(-0.5, 0.5)])
        g = orc.cost_grad_f64(gL, G1, gR, G2, x, y, s, P)
        c = orc.cost_grad_f64(cL, G1, cR, G2, x, y, s, P)
        assert c[0] == pytest.approx(g[0], rel=1e-12)
        np.testing.assert_allclose(c[1], g[1], rtol=1e-9, atol=1e-12)
        for cg, gg in ((c[2], g[2]), (c[4], g[4])):  # dI1, dI2
            np.testing.assert_allclose(cg[..., :3].sum(-1), gg, rtol=1e-9, atol=1e-12)
            assert not cg[..., 3].any()
        for cg, gg in ((c[3], g[3]), (c[5], g[5])):  # dG1, dG2
            np.testing.assert_allclose(cg, gg, rtol=1e-9, atol=1e-12)
