"""Oracle pins for PMHuberStereo::depthMapInpaint's kernels.

occmap      -- brute-force L/R test (the consistency test with a 5 degree
               normal tolerance), away from the thresholds; layout (v, v, v, 1).
normal_depth-- equals fill_state on fields with no invalid pixel (fill_state
               converts the same way when it has nothing to fill).
pd_step_inpaint -- FP64 restatement of one step (per-channel dual projection,
               lambda scaled by the mask); a constant field is a fixed point;
               with the projection inactive and mask 1 it equals the
               depthMap smoother's step bit for bit.
depthmap_inpaint -- runs the reference sequence and recovers the scene.
"""
import numpy as np
import pytest

from test_oracle_pipeline import lr_fields


@pytest.mark.parametrize("lr", [1, 0])
def test_occmap_brute_force(orc, lr):
    l, r = lr_fields(orc)
    p, q = (l, r) if lr else (r, l)
    H, W = p.shape[:2]
    got = orc.occmap(p, q, 0.5, lr)
    assert np.all(got[..., 3] == 1.0)
    assert np.array_equal(got[..., 0], got[..., 1]) and np.array_equal(got[..., 0], got[..., 2])
    for y in range(H):
        for x in range(W):
            xp = orc.state_transform(p[y, x], x, y)
            sr = np.full(4, 1e5, np.float32) if not (0 <= xp < W) else \
                q[y, min(int(np.floor(xp + 0.5)), W - 1)]
            if lr:
                n1, n2 = orc.state_to_normal(p[y, x]), orc.state_to_normal(orc.state_viewprop(sr))
            else:
                n1, n2 = orc.state_to_normal(sr), orc.state_to_normal(orc.state_viewprop(p[y, x]))
            nd = float(np.dot(n1.astype(np.float64), n2.astype(np.float64)))
            dd = abs(x - orc.state_transform(sr, xp, y))
            ang = np.degrees(np.arccos(np.clip(nd, -1, 1)))
            if abs(dd - 0.5) < 1e-4 or abs(ang - 5.0) < 1e-3:
                continue
            assert got[y, x, 0] == (1.0 if (dd < 0.5 and ang < 5.0) else 0.0), (x, y)
    assert 0.02 < got[..., 0].mean() < 0.98


@pytest.mark.parametrize("lr", [1, 0])
def test_normal_depth_equals_fill_state_on_valid_fields(orc, lr):
    l, r = lr_fields(orc)
    st = l if lr else r
    assert (np.sqrt((st.astype(np.float64) ** 2).sum(-1)) >= 0.1).all()
    np.testing.assert_array_equal(orc.normal_depth(st, 8.0, lr), orc.fill_state(st, 8.0, lr))


def test_pd_step_inpaint_fp64_and_fixed_point(orc):
    rng = np.random.default_rng(18)
    H, W = 13, 17
    w = (0.1 + rng.random((H, W))).astype(np.float32)
    mask = np.zeros((H, W, 4), np.float32)
    mask[..., 0] = (rng.random((H, W)) < 0.7).astype(np.float32)
    mask[..., 3] = 1
    # constant field: fixed point whatever the mask
    f = np.zeros((H, W, 4), np.float32)
    f[...] = [0.2, -0.1, 0.4, 1.0]
    u, p = orc.pd_step_inpaint(f, np.zeros((H, W, 8), np.float32), f, w, mask, 5.0)
    np.testing.assert_allclose(u[..., :3], f[..., :3], atol=1e-7)
    assert not p.any()
    # FP64 restatement of PDHuberWeightedInpaint.cl:48-103
    last = rng.standard_normal((H, W, 4)).astype(np.float32)
    pl = (0.3 * rng.standard_normal((H, W, 8))).astype(np.float32)
    pl[..., 3] = pl[..., 7] = 0
    fr = rng.random((H, W, 4)).astype(np.float32)
    lam = 7.0
    u, p = orc.pd_step_inpaint(last, pl, fr, w, mask, lam)
    sig = tau = 1 / np.sqrt(8)
    L = last.astype(np.float64)[..., :3]
    Wd = w.astype(np.float64)

    def dual(gx, gy):
        if gx < 0 or gy < 0:
            return np.zeros(3), np.zeros(3)
        x1, y1 = min(gx + 1, W - 1), min(gy + 1, H - 1)
        wt = Wd[gy, gx]
        lo = (pl[gy, gx, :3] + sig * wt * (L[gy, gx] - L[gy, x1])) / (1 + sig * 0.001 / wt)
        hi = (pl[gy, gx, 4:7] + sig * wt * (L[gy, gx] - L[y1, gx])) / (1 + sig * 0.001 / wt)
        s = np.maximum(1.0, np.sqrt(lo * lo + hi * hi))   # every channel alone
        return lo / s, hi / s

    for y in range(H):
        for x in range(W):
            lo, hi = dual(x, y)
            lol, _ = dual(x - 1, y)
            _, hiu = dual(x, y - 1)
            div = Wd[y, x] * (lo - lol + hi - hiu)
            lm = lam * mask[y, x, 0]
            un = (L[y, x] + tau * (lm * fr[y, x, :3] - div)) / (1 + tau * lm)
            un = un + 0.5 * (un - L[y, x])
            np.testing.assert_allclose(u[y, x, :3], un, rtol=1e-5, atol=1e-5)
            np.testing.assert_allclose(p[y, x, :3], lo, rtol=1e-5, atol=1e-6)
            assert u[y, x, 3] == 1.0


def test_pd_step_inpaint_equals_plain_step_when_projection_inactive(orc):
    rng = np.random.default_rng(19)
    H, W = 11, 15
    w = (0.5 + 0.5 * rng.random((H, W))).astype(np.float32)
    last = (0.01 * rng.standard_normal((H, W, 4))).astype(np.float32)
    last[..., 3] = 1
    pl = (0.01 * rng.standard_normal((H, W, 8))).astype(np.float32)
    pl[..., 3] = pl[..., 7] = 0
    f = (0.01 * rng.standard_normal((H, W, 4))).astype(np.float32)
    ones = np.ones((H, W, 4), np.float32)
    a = orc.pd_step_inpaint(last, pl, f, w, ones, 5.0)
    b = orc.pd_step(last, pl, f, w, 5.0)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_depthmap_inpaint_recovers_scene(orc):
    from cvt_amd import synth
    pr = synth.make_pair(96, 64, 4, 16.0, 0.01, seed=3)
    r = orc.depthmap_inpaint(pr.left, pr.right, 5, 16.0, 5, pd_iters=20, want_states=True)
    assert r["dmap"].shape == (64, 96)
    assert synth.bad_pixels(r["dmap"], pr) < 10.0
