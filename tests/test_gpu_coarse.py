"""Coarse-to-fine start: the 2x2 image reduction and the plane hand-up, bit-exact
against numpy's float32 arithmetic in the same order, the hand-up's geometry
against its closed form, and the batched coarse-to-fine run against single
pairs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available()
    from cvt_amd import ops
    return ops


def T(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("W,H", [(64, 48), (77, 51), (1, 1), (2, 3)])
def test_pyr_down_bitexact(dev, W, H):
    rng = np.random.default_rng(W * 100 + H)
    img = rng.random((H, W), dtype=np.float32)
    got = dev.pyr_down(T(img)).cpu().numpy()
    w, h = (W + 1) // 2, (H + 1) // 2
    x0 = 2 * np.arange(w)
    x1 = np.minimum(x0 + 1, W - 1)
    y0 = 2 * np.arange(h)
    y1 = np.minimum(y0 + 1, H - 1)
    a, b = img[y0][:, x0], img[y0][:, x1]
    c, d = img[y1][:, x0], img[y1][:, x1]
    exp = np.float32(0.25) * ((a + b) + (c + d))
    assert got.shape == (h, w)
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("W,H", [(64, 48), (77, 51)])
def test_plane_up_bitexact_and_geometry(dev, W, H):
    rng = np.random.default_rng(7)
    w, h = (W + 1) // 2, (H + 1) // 2
    st = np.zeros((h, w, 4), np.float32)
    st[..., 0] = rng.uniform(0.7, 1.3, (h, w))
    st[..., 1] = rng.uniform(-0.3, 0.3, (h, w))
    st[..., 2] = rng.uniform(-40, 40, (h, w))
    st[..., 3] = 5.0
    got = dev.plane_up(T(st), W, H).cpu().numpy()
    ys, xs = np.mgrid[0:H, 0:W]
    src = st[np.minimum(ys // 2, h - 1), np.minimum(xs // 2, w - 1)]
    a, b, c = src[..., 0], src[..., 1], src[..., 2]
    exp_c = np.float32(2.0) * c + np.float32(0.5) * ((np.float32(1.0) - a) - b)
    assert np.array_equal(got[..., 0], a) and np.array_equal(got[..., 1], b)
    assert np.array_equal(got[..., 2], exp_c)
    assert not got[..., 3].any()
    # geometry: the fine-level disparity at a pixel centre is twice the coarse
    # disparity at the matching coarse coordinate (x_c = (x_f - 1/2) / 2)
    a64, b64, c64 = (src[..., k].astype(np.float64) for k in range(3))
    xc, yc = (xs - 0.5) / 2, (ys - 0.5) / 2
    d_coarse = xc - (a64 * xc + b64 * yc + c64)
    g = got.astype(np.float64)
    d_fine = xs - (g[..., 0] * xs + g[..., 1] * ys + g[..., 2])
    assert np.allclose(d_fine, 2 * d_coarse, atol=1e-3)


def test_coarse_batch_equals_single_and_converges(dev):
    import torch
    from cvt_amd import _lib
    from cvt_amd.synth import bad_pixels, make_pair
    prs = [make_pair(W, H, n, d, 0.01, seed=60 + k)
           for k, (W, H, n, d) in enumerate([(200, 120, 4, 32.0), (153, 97, 3, 24.0)])]
    pairs = [(T(p.left), T(p.right), p.dmax, 1.0) for p in prs]
    opts = _lib.default_options(sweep=3, maxdispdiff=1.0, pd_iters=40, coarse_levels=2,
                                coarse_iters=3, grad_steps=1)
    got = dev.batch_depthmap(pairs, 5, 2, opts, states=True)
    for (l, r, d, s), g, p in zip(pairs, got, prs):
        ref = dev.depthmap(l, r, 5, d, 2, 4, s, opts, states=True)
        assert torch.equal(ref["dmap"], g["dmap"])
        assert torch.equal(ref["left_states"], g["left_states"])
        bad = bad_pixels(g["dmap"].cpu().numpy(), p)
        print(f"{p.left.shape}: bad-1.0px {bad:.2f}%")
        assert bad < 10.0


def test_coarse_joint_views(dev):
    """Both views of a coarse level in one launch per sweep (each reading the
    other view's scatter of the previous iteration): the same bits batched as
    one pair at a time, a schedule of its own at the coarse levels (the left
    field differs from the reference order's), no effect without coarse
    levels (full resolution keeps the reference's view order), and as
    accurate."""
    import torch
    from cvt_amd import _lib
    from cvt_amd.synth import bad_pixels, make_pair
    prs = [make_pair(W, H, n, d, 0.01, seed=70 + k)
           for k, (W, H, n, d) in enumerate([(200, 120, 4, 32.0), (141, 89, 3, 24.0)])]
    pairs = [(T(p.left), T(p.right), p.dmax, 1.0) for p in prs]
    kw = dict(sweep=3, maxdispdiff=1.0, pd_iters=40, coarse_levels=2, coarse_iters=3,
              grad_steps=1)
    joint = _lib.default_options(coarse_joint_views=1, **kw)
    seq = _lib.default_options(**kw)
    got = dev.batch_depthmap(pairs, 5, 2, joint, states=True)
    for (l, r, d, s), g, p in zip(pairs, got, prs):
        ref = dev.depthmap(l, r, 5, d, 2, 4, s, joint, states=True)
        assert torch.equal(ref["dmap"], g["dmap"])
        assert torch.equal(ref["left_states"], g["left_states"])
        bad = bad_pixels(g["dmap"].cpu().numpy(), p)
        print(f"{p.left.shape}: joint views, bad-1.0px {bad:.2f}%")
        assert bad < 10.0
    other = dev.batch_depthmap(pairs, 5, 2, seq, states=True)
    assert any(not torch.equal(a["left_states"], b["left_states"]) for a, b in zip(got, other))
    flat = dict(sweep=3, maxdispdiff=1.0, pd_iters=20, grad_steps=1)
    a = dev.batch_depthmap(pairs, 5, 2, _lib.default_options(coarse_joint_views=1, **flat),
                           states=True)
    b = dev.batch_depthmap(pairs, 5, 2, _lib.default_options(**flat), states=True)
    for x, y in zip(a, b):
        assert torch.equal(x["dmap"], y["dmap"])
        assert torch.equal(x["left_states"], y["left_states"])
        assert torch.equal(x["right_states"], y["right_states"])
    with pytest.raises(_lib.PmhError, match="coarse_joint_views"):
        dev.batch_depthmap(pairs, 5, 1, _lib.default_options(coarse_joint_views=2))


def _pyr_down(img):
    """The 2x2 box reduction in float32, in the kernel's order (pinned bit for
    bit by test_pyr_down_bitexact)."""
    H, W = img.shape
    w, h = (W + 1) // 2, (H + 1) // 2
    x0 = 2 * np.arange(w)
    x1 = np.minimum(x0 + 1, W - 1)
    y0 = 2 * np.arange(h)
    y1 = np.minimum(y0 + 1, H - 1)
    a, b = img[y0][:, x0], img[y0][:, x1]
    c, d = img[y1][:, x0], img[y1][:, x1]
    return np.float32(0.25) * ((a + b) + (c + d))


def _plane_up(st, W, H):
    """The plane hand-up in float32 (pinned by test_plane_up_bitexact_and_geometry)."""
    h, w = st.shape[:2]
    ys, xs = np.mgrid[0:H, 0:W]
    src = st[np.minimum(ys // 2, h - 1), np.minimum(xs // 2, w - 1)]
    out = np.zeros((H, W, 4), np.float32)
    out[..., 0], out[..., 1] = src[..., 0], src[..., 1]
    out[..., 2] = (np.float32(2.0) * src[..., 2] +
                   np.float32(0.5) * ((np.float32(1.0) - src[..., 0]) - src[..., 1]))
    return out


def _compose_level(orc, left, right, P, dmax, iterations, maxdisp, maxdeg, pd_iters, joint,
                   grad_steps=0, radius=0.5):
    """One coarse level from random planes: the depthMap sequence
    (PMHuberStereo.cpp:90-266) over the oracle's pinned per-kernel functions,
    red-black.  joint: both views propagate from the other's scatter of the
    previous iteration (the left view's buffer starts cleared too); otherwise
    the reference's order.  grad_steps > 0: each view's propagation is
    followed by that many gradient refinement steps (orc_grad_refine, the
    iteration's theta and smooth prior) before its scatter, as at every level
    of the product.  The smoothing after the last iteration feeds nothing and
    is left out.  Returns the final left and right fields."""
    from test_oracle_compose import as_viewin, pdrof_pingpong, scatter, smoothstep
    H, W = left.shape
    f32 = np.float32
    lg, rg = orc.gradxy(left), orc.gradxy(right)
    lw, rw = orc.weight(left), orc.weight(right)
    lsm = np.zeros((H, W, 4), np.float32)
    rsm = np.zeros((H, W, 4), np.float32)
    m1 = [orc.init(left, lg, right, rg, P, dmax, 1), None]
    m2 = [orc.init(right, rg, left, lg, P, dmax, 0), None]
    empty = (np.zeros((H, W), np.int32), np.zeros((H, W, 4, 4), np.float32))
    vb1, vb2 = empty, empty
    theta = f32(0)

    def prop(old, I1, G1, I2, G2, sm, lr, it, view):
        vb, src = as_viewin(*view)
        out = old.copy()
        for s in (1, 2):
            out = orc.propagate(old, I1, G1, I2, G2, sm, float(theta), P, dmax, lr, it,
                                viewin=vb, viewsrc=src, sweep=s, out=out)
        if grad_steps:
            out = orc.grad_refine(out, I1, G1, I2, G2, sm, float(theta), P, dmax, lr,
                                  grad_steps, radius)
        return out

    for it in range(iterations):
        swap = it & 1
        m1[1 - swap] = prop(m1[swap], left, lg, right, rg, lsm, 1, it, vb2)
        if joint:
            m2[1 - swap] = prop(m2[swap], right, rg, left, lg, rsm, 0, it, vb1)
            vb1 = scatter(orc, m1[1 - swap])
        else:
            vb1 = scatter(orc, m1[1 - swap])
            m2[1 - swap] = prop(m2[swap], right, rg, left, lg, rsm, 0, it, vb1)
        vb2 = scatter(orc, m2[1 - swap])
        if it == iterations - 1:
            break
        lam = float(f32(f32(theta * f32(50)) + f32(5)))
        tmp = orc.fill_state(orc.consistency(m1[1 - swap], m2[1 - swap], maxdisp, maxdeg, 1),
                             dmax, 1)
        lsm = pdrof_pingpong(orc, lsm, tmp, lw, lam, pd_iters)
        tmp = orc.fill_state(orc.consistency(m2[1 - swap], m1[1 - swap], maxdisp, maxdeg, 0),
                             dmax, 0)
        rsm = pdrof_pingpong(orc, rsm, tmp, rw, lam, pd_iters)
        if it >= 5:
            theta = smoothstep(f32(f32(f32(it) - f32(5)) / f32(f32(iterations) - f32(5))))
    fin = 1 - ((iterations - 1) & 1)
    return m1[fin], m2[fin]


@pytest.mark.parametrize("joint", [0, 1])
def test_coarse_path_matches_oracle_composition(dev, orc, joint):
    """The coarse-to-fine path against an independent composition from pinned
    pieces.  With no full-resolution iterations, pmh_batch_depthmap returns
    the L/R check and fill of the handed-up planes, their costs recomputed at
    full resolution; the composition runs the 2x2 reduction and the hand-up in
    numpy, the coarse level over the oracle's per-kernel functions (in the
    reference's view order, or both views per launch), then the oracle's
    window cost, consistency and fill_depthmap.  Disparity within 0.5 px on
    at least 99% of the pixels."""
    from cvt_amd import _lib
    from cvt_amd.synth import make_pair
    W, H, P, Pc, dmax, iters, pdi = 150, 92, 3, 2, 16.0, 3, 10
    pr = make_pair(W, H, 4, dmax, 0.01, seed=90 + joint)
    opts = _lib.default_options(sweep=3, maxdispdiff=1.0, coarse_levels=1, coarse_iters=iters,
                                coarse_patch=Pc, coarse_pd_iters=pdi,
                                coarse_joint_views=joint)
    got = dev.batch_depthmap([(T(pr.left), T(pr.right), dmax, 1.0)], P, 0, opts)[0]["dmap"]
    got = got.cpu().numpy()
    L1, R1 = _pyr_down(pr.left), _pyr_down(pr.right)
    mL, mR = _compose_level(orc, L1, R1, Pc, 0.5 * dmax, iters, 1.0, 10.0, pdi, joint)
    upL, upR = _plane_up(mL, W, H), _plane_up(mR, W, H)
    lg, rg = orc.gradxy(pr.left), orc.gradxy(pr.right)
    upL[..., 3] = orc.cost_field(pr.left, lg, pr.right, rg, upL, P)
    upR[..., 3] = orc.cost_field(pr.right, rg, pr.left, lg, upR, P)
    cons = orc.consistency(upL, upR, 1.0, 10.0, 1)
    exp = orc.fill_depthmap(cons, 1.0)
    frac = float(np.mean(np.abs(got - exp) <= 0.5))
    print(f"coarse path (joint={joint}) within 0.5 px of the oracle composition: {frac:.4f}")
    assert frac >= 0.99


def test_batched_method_matches_oracle_composition(dev, orc):
    """The batched method end to end against an oracle composition: a coarse
    level with both views per launch and two gradient steps per view and
    iteration, the hand-up, then one
    full-resolution red-black iteration without randomized candidates, each
    view's propagation followed by two gradient refinement steps
    (orc_grad_refine) before its scatter, and the final L/R check and fill.
    Disparity within 0.5 px on at least 99% of the pixels."""
    from test_oracle_compose import as_viewin, scatter
    from cvt_amd import _lib
    from cvt_amd.synth import bad_pixels, make_pair
    W, H, P, Pc, dmax, iters, pdi = 150, 92, 3, 2, 16.0, 3, 10
    pr = make_pair(W, H, 4, dmax, 0.01, seed=95)
    opts = _lib.default_options(sweep=3, maxdispdiff=1.0, coarse_levels=1, coarse_iters=iters,
                                coarse_patch=Pc, coarse_pd_iters=pdi, coarse_joint_views=1,
                                coarse_random_refine=2, coarse_random_neighbours=6,
                                grad_steps=2, random_refine=0, random_neighbours=0)
    got = dev.batch_depthmap([(T(pr.left), T(pr.right), dmax, 1.0)], P, 1, opts)[0]["dmap"]
    got = got.cpu().numpy()
    L1, R1 = _pyr_down(pr.left), _pyr_down(pr.right)
    mL, mR = _compose_level(orc, L1, R1, Pc, 0.5 * dmax, iters, 1.0, 10.0, pdi, 1,
                            grad_steps=2)
    left, right = pr.left, pr.right
    lg, rg = orc.gradxy(left), orc.gradxy(right)
    m1 = [_plane_up(mL, W, H), None]
    m2 = [_plane_up(mR, W, H), None]
    m1[0][..., 3] = orc.cost_field(left, lg, right, rg, m1[0], P)
    m2[0][..., 3] = orc.cost_field(right, rg, left, lg, m2[0], P)
    zero = np.zeros((H, W, 4), np.float32)

    def prop(old, I1, G1, I2, G2, lr, view):
        vb, src = as_viewin(*view)
        out = old.copy()
        for s in (1, 2):
            out = orc.propagate(old, I1, G1, I2, G2, zero, 0.0, P, dmax, lr, 0, viewin=vb,
                                viewsrc=src, sweep=s, out=out, random_neighbours=0,
                                random_refine=0)
        return orc.grad_refine(out, I1, G1, I2, G2, zero, 0.0, P, dmax, lr, 2, 0.5)

    empty = (np.zeros((H, W), np.int32), np.zeros((H, W, 4, 4), np.float32))
    m1[1] = prop(m1[0], left, lg, right, rg, 1, empty)
    m2[1] = prop(m2[0], right, rg, left, lg, 0, scatter(orc, m1[1]))
    exp = orc.fill_depthmap(orc.consistency(m1[1], m2[1], 1.0, 10.0, 1), 1.0)
    frac = float(np.mean(np.abs(got - exp) <= 0.5))
    print(f"batched method within 0.5 px of the oracle composition: {frac:.4f}; bad-1.0px "
          f"GPU {bad_pixels(got, pr):.2f}%, oracle {bad_pixels(exp, pr):.2f}%")
    assert frac >= 0.99
