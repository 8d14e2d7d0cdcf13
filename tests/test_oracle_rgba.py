"""RGBA inputs of the oracle (CL_RGBA images: colour sums over R, G, B).

Pins: an RGBA image whose R = G = B equal the gray image gives the gray
results bit for bit (gradients, weights, costs, init, propagation, depthMap);
per-channel constant colour offsets give the closed-form colour term
0.05 * sum_c min(|d_c|, 0.04); the support weight of a tap is
exp(-(|dR| + |dG| + |dB|) F), checked by brute force in FP64 on a window with
one colour step."""
import numpy as np
import pytest

P = 3


def rgba_of(gray, alpha=1.0):
    return np.stack([gray, gray, gray, np.full_like(gray, alpha)], -1).astype(np.float32)


def test_gray_replicated_rgba_is_bitexact(orc):
    rng = np.random.default_rng(11)
    H, W = 24, 40
    L = rng.random((H, W)).astype(np.float32)
    R = np.roll(L, -3, axis=1) + 0.01 * rng.standard_normal((H, W)).astype(np.float32)
    R = R.astype(np.float32)
    Lc, Rc = rgba_of(L), rgba_of(R, alpha=0.3)
    assert np.array_equal(orc.gradxy(L), orc.gradxy(Lc))
    assert np.array_equal(orc.weight(R), orc.weight(Rc))
    GL, GR = orc.gradxy(L), orc.gradxy(R)
    st = orc.init(L, GL, R, GR, P, 8.0, 1)
    assert np.array_equal(st, orc.init(Lc, GL, Rc, GR, P, 8.0, 1))
    sm = np.zeros((H, W, 4), np.float32)
    a = orc.propagate(st, L, GL, R, GR, sm, 0.0, P, 8.0, 1, 0)
    b = orc.propagate(st, Lc, GL, Rc, GR, sm, 0.0, P, 8.0, 1, 0)
    assert np.array_equal(a, b)
    d1 = orc.depthmap(L, R, P, 8.0, 2, 1.0, pd_iters=10, sweep=1)["dmap"]
    d2 = orc.depthmap(Lc, Rc, P, 8.0, 2, 1.0, pd_iters=10, sweep=1)["dmap"]
    assert np.array_equal(d1, d2)


@pytest.mark.parametrize("off", [(0.01, 0.02, 0.0), (0.05, -0.01, 0.03), (-0.2, 0.1, 0.035)])
def test_constant_colour_offsets(orc, off):
    H, W = 30, 40
    base = np.array([0.4, 0.5, 0.3, 1.0], np.float32)
    I1 = np.broadcast_to(base, (H, W, 4)).astype(np.float32).copy()
    I2 = I1.copy()
    I2[..., :3] += np.array(off, np.float32)
    G = np.zeros((H, W, 4), np.float32)   # equal gradient images: only the colour term
    s = np.array([1.0, 0.0, 0.0, 0.0], np.float32)
    c = orc.cost(I1, G, I2, G, 20, 15, s, P)
    d = (I2 - I1)[0, 0, :3].astype(np.float64)
    assert c == pytest.approx(0.05 * np.minimum(np.abs(d), 0.04).sum(), rel=2e-5)


def test_colour_support_weights_fp64(orc):
    rng = np.random.default_rng(3)
    H, W = 20, 30
    I1 = (0.3 + 0.4 * rng.random((H, W, 4))).astype(np.float32)
    I2 = (I1 + np.float32(0.02) * rng.standard_normal((H, W, 4))).astype(np.float32)
    G1 = (0.01 * rng.standard_normal((H, W, 4))).astype(np.float32)
    G2 = (0.01 * rng.standard_normal((H, W, 4))).astype(np.float32)
    s = np.array([1.0, 0.0, 0.0, 0.0], np.float32)   # x' = x: no interpolation
    x, y = 15, 10
    d1, d2, g1, g2 = (a.astype(np.float64) for a in (I1, I2, G1, G2))
    num = den = 0.0
    for qy in range(y - P, y + P + 1):
        for qx in range(x - P, x + P + 1):
            C = 0.05 * np.minimum(np.abs(d1[qy, qx, :3] - d2[qy, qx, :3]), 0.04).sum() + \
                0.95 * np.minimum(np.abs(g1[qy, qx] - g2[qy, qx]), 0.01).sum()
            t = min(np.hypot(qx - x, qy - y) / 26.0, 1.0)
            F = t * t * (3 - 2 * t) * 39 + 5
            w = np.exp(-np.abs(d1[y, x, :3] - d1[qy, qx, :3]).sum() * F)
            num += w * C
            den += w
    assert orc.cost(I1, G1, I2, G2, x, y, s, P) == pytest.approx(num / den, rel=1e-5)
