"""Pins of the oracle's gradient plane refinement (orc_grad_refine; the
product's directed refinement, no reference counterpart).

  - one step moves x' by exactly the radius at the window's farthest tap
    (|de| + P (|da| + |db|) = r, the step's defining constraint): a wrong sign
    or index in the (e, a, b) -> (a, b, c) conversion moves x' elsewhere;
  - a moved plane's x' at the pixel itself moves against the cost's
    derivative there (descent), checked with the FP64 gradient;
  - the stored cost is the window cost of the stored plane, never above the
    input's, and planes do move;
  - zero steps and planes without a finite cost change nothing.
"""
import numpy as np
import pytest

f32 = np.float32


@pytest.fixture(scope="module")
def field(orc):
    from cvt_amd.synth import make_pair
    pr = make_pair(64, 40, 3, 12.0, 0.01, seed=11)
    P, dmax = 3, 12.0
    res = orc.depthmap(pr.left, pr.right, P, dmax, 2, 1.0, want_states=True, pd_iters=20)
    L, R = pr.left, pr.right
    return L, R, orc.gradxy(L), orc.gradxy(R), res["left_states"], P, dmax


def test_one_step_moves_x_by_the_radius(orc, field):
    L, R, GL, GR, st, P, dmax = field
    H, W = L.shape
    sm = np.zeros((H, W, 4), np.float32)
    r = 0.25
    out = orc.grad_refine(st, L, GL, R, GR, sm, 0.0, P, dmax, 1, 1, r)
    moved = np.any(out[..., :3] != st[..., :3], axis=-1)
    assert moved.mean() > 0.2
    ys, xs = np.nonzero(moved)
    for y, x in zip(ys[:200], xs[:200]):
        a0, b0, c0 = (float(v) for v in st[y, x, :3])
        a1, b1, c1 = (float(v) for v in out[y, x, :3])
        # x' displacement over the window's taps
        dmax_win = max(abs((a1 - a0) * qx + (b1 - b0) * qy + (c1 - c0))
                       for qx in (x - P, x + P) for qy in (y - P, y + P))
        assert abs(dmax_win - r) < 2e-4 * (1 + abs(c0)), (y, x, dmax_win)
        # descent: x' at the pixel moves against d cost / d x'(x, y)
        de = (a1 - a0) * x + (b1 - b0) * y + (c1 - c0)
        g = orc.cost_grad_f64(L, GL, R, GR, x, y, st[y, x, :3], P, images=False)[1]
        ge = g[2]  # d cost / d c = d cost / d x'(x, y)
        if abs(ge) > 1e-6 and abs(de) > 1e-6:
            assert np.sign(de) == -np.sign(ge), (y, x, de, ge)


def test_cost_is_stored_and_never_rises(orc, field):
    L, R, GL, GR, st, P, dmax = field
    H, W = L.shape
    rng = np.random.default_rng(5)
    sm = np.zeros((H, W, 4), np.float32)
    sm[..., 0:2] = rng.normal(0, 0.1, (H, W, 2))
    sm[..., 2] = rng.uniform(0.2, 0.6, (H, W))
    for theta in (0.0, 0.5):
        out = orc.grad_refine(st, L, GL, R, GR, sm, theta, P, dmax, 1, 2, 0.5)
        assert np.array_equal(out[..., 3], orc.cost_field(L, GL, R, GR, out, P))
        if theta == 0.0:
            assert np.all(out[..., 3] <= st[..., 3])
    assert np.array_equal(orc.grad_refine(st, L, GL, R, GR, sm, 0.0, P, dmax, 1, 0, 0.5), st)


def test_planes_without_a_cost_are_kept(orc, field):
    L, R, GL, GR, st, P, dmax = field
    H, W = L.shape
    bad = st.copy()
    bad[::3, ::5, 0] = np.nan
    bad[1::3, ::7, 0] = 0.0  # x' = b y + 1e30: outside the image at every tap
    bad[1::3, ::7, 2] = 1e30
    bad[1::3, ::7, 3] = 1e5
    out = orc.grad_refine(bad, L, GL, R, GR, np.zeros((H, W, 4), np.float32), 0.0, P, dmax, 1,
                          2, 0.5)
    keep = ~np.isfinite(bad[..., 0]) | (bad[..., 2] == 1e30)
    assert np.array_equal(out[keep], bad[keep], equal_nan=True)
