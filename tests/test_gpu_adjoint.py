"""Adjoint of the window cost (HIP, through the C-ABI) against the oracle's FP64
gradient and central finite differences; the gradxy adjoint by the dot-product
identity; the autograd pair; the gradient-step plane refinement.

Plane gradients are compared in local coordinates at the pixel, (d/de, d/da,
d/db) with x' = e + a (qx - x) + b (qy - y), which are O(1) and free of the
x * d/dc cancellation of the raw (a, b, c) components.

The cost is piecewise linear in the plane: its derivative jumps where a tap's
x' crosses a bilinear cell edge or a truncation threshold (min(|d|, t)) or
where |d| crosses 0.  With 1225 taps per window some tap is almost always
within float rounding (GPU, FP32) of such a kink, so a few windows pick the
other side of one kink and their gradient differs by one tap's jump
(~1e-3..1e-2 relative).  The bars are therefore statistical: the median
agreement is at rounding level and only a small fraction of pixels may carry
a jump; finite differences (step 1e-3 px of x') likewise cross kinks in some
taps, so they are held to 1e-3 relative for most components and 1e-2 for
nearly all.
"""
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


def N(t):
    return t.detach().cpu().numpy()


def local(d, x, y):
    """(dc, da, db) of a plane gradient -> (d/de, d/da_local, d/db_local)."""
    return np.array([d[2], d[0] - x * d[2], d[1] - y * d[2]])


@pytest.fixture(scope="module")
def big(dev, orc):
    from cvt_amd import _lib, synth
    pr = synth.pair_1242x375()
    L, R = T(pr.left), T(pr.right)
    res = dev.depthmap(L, R, 17, 192.0, 2, 4, 1.0,
                       _lib.default_options(sweep=_lib.SWEEP_REDBLACK, maxdispdiff=1.0),
                       states=True)
    st = N(res["left_states"]).copy()
    rng = np.random.default_rng(0)
    st[..., 2] += rng.uniform(-0.7, 0.7, st.shape[:2]).astype(np.float32)  # off the optimum
    return pr, st


def test_plane_gradient_1024_pixels_vs_fp64_and_fd(dev, orc, big):
    pr, st = big
    P = 17
    L, R = pr.left, pr.right
    GL, GR = orc.gradxy(L), orc.gradxy(R)
    H, W = L.shape
    g = N(dev.cost_backward(T(L), T(GL), T(R), T(GR), T(st), T(np.ones((H, W), np.float32)), P,
                            images=False)[0])
    rng = np.random.default_rng(1024)
    xs, ys = rng.integers(0, W, 1024), rng.integers(0, H, 1024)
    errs, fd_errs = [], []
    for x, y in zip(xs, ys):
        x, y = int(x), int(y)
        s = st[y, x, :3].astype(np.float64)
        c, dp, *_ = orc.cost_grad_f64(L, GL, R, GR, x, y, s, P, images=False)
        if c >= 1e5:
            assert not g[y, x, :3].any()
            errs.append(0.0)
            continue
        ref = local(dp, x, y)
        got = local(g[y, x, :3].astype(np.float64), x, y)
        scale = np.abs(ref).max() + 1e-6
        errs.append(np.abs(got - ref).max() / scale)
        # central differences in FP64, step 1e-3 px of x' at the window edge
        h = np.array([1e-3, 1e-3 / P, 1e-3 / P])
        for j in range(3):
            def shifted(t):
                e = np.zeros(3)
                e[j] = t
                # local -> (a, b, c): a += e_a, b += e_b, c += e_e - e_a x - e_b y
                return s + np.array([e[1], e[2], e[0] - e[1] * x - e[2] * y])
            f = lambda t: orc.cost_f64(L, GL, R, GR, x, y, shifted(t), P)
            fd1 = (f(h[j]) - f(-h[j])) / (2 * h[j])
            fd2 = (f(h[j] / 2) - f(-h[j] / 2)) / h[j]
            if abs(fd1 - fd2) > 1e-3 * (abs(fd1) + 1e-5):
                continue  # the FD itself is not converged: a kink near the centre
            fd_errs.append(abs(fd1 - ref[j]) / scale)
    errs, fd_errs = np.array(errs), np.array(fd_errs)
    print(f"GPU vs FP64 (1024 px): median {np.median(errs):.2e}, "
          f"{np.mean(errs < 2e-3) * 100:.1f}% < 2e-3, p99 {np.quantile(errs, 0.99):.2e}; "
          f"FD vs FP64 ({len(fd_errs)} comps): {np.mean(fd_errs < 1e-3) * 100:.1f}% < 1e-3, "
          f"{np.mean(fd_errs < 1e-2) * 100:.1f}% < 1e-2")
    assert np.median(errs) < 1e-4
    assert np.mean(errs < 2e-3) >= 0.97
    assert np.quantile(errs, 0.99) < 5e-2
    assert len(fd_errs) > 1000
    assert np.mean(fd_errs < 1e-3) >= 0.90 and np.mean(fd_errs < 1e-2) >= 0.99


@pytest.mark.parametrize("colour", [False, True], ids=["gray", "rgba"])
def test_image_gradients_vs_fp64(dev, orc, colour):
    """Plane and image gradients of a propagated field against the oracle's
    FP64 gradient; RGBA: three colour channels each (alpha: none)."""
    from cvt_amd.synth import make_pair
    pr = make_pair(80, 48, 3, 16.0, 0.01, seed=9, colour=colour)
    P = 5
    L, R = pr.left, pr.right
    GL, GR = orc.gradxy(L), orc.gradxy(R)
    H, W = L.shape[:2]
    st = orc.init(L, GL, R, GR, P, 16.0, 1)
    sm = np.zeros((H, W, 4), np.float32)
    for it in range(3):
        st = orc.propagate(st, L, GL, R, GR, sm, 0.0, P, 16.0, 1, it)
    rng = np.random.default_rng(3)
    gout = rng.standard_normal((H, W)).astype(np.float32)
    dst, dI1, dG1, dI2, dG2 = [N(t) for t in dev.cost_backward(T(L), T(GL), T(R), T(GR), T(st),
                                                                T(gout), P)]
    rp, rI1, rG1, rI2, rG2 = orc.cost_grad_field_f64(L, GL, R, GR, st, P, gout)
    assert dI1.shape == L.shape and dI2.shape == L.shape
    if colour:
        assert not dI1[..., 3].any() and not dI2[..., 3].any()
    for got, ref, name in ((dI1, rI1, "dI1"), (dG1, rG1, "dG1"), (dI2, rI2, "dI2"),
                           (dG2, rG2, "dG2")):
        scale = np.abs(ref).max()
        bad = np.abs(got - ref) > 2e-3 * scale + 1e-3 * np.abs(ref)
        assert bad.mean() < 0.003, (name, bad.mean(), scale)
    loc_g = np.stack([dst[..., 2], dst[..., 0] - np.arange(W) * dst[..., 2],
                      dst[..., 1] - np.arange(H)[:, None] * dst[..., 2]], -1)
    loc_r = np.stack([rp[..., 2], rp[..., 0] - np.arange(W) * rp[..., 2],
                      rp[..., 1] - np.arange(H)[:, None] * rp[..., 2]], -1)
    err = np.abs(loc_g - loc_r).max(-1) / (np.abs(loc_r).max(-1) + 1e-6)
    print(f"plane gradient field: median {np.median(err):.2e}, {np.mean(err < 2e-3) * 100:.1f}% "
          f"< 2e-3, p99 {np.quantile(err, 0.99):.2e}")
    # a pixel off the FP64 value by more than rounding has a tap within
    # rounding of a kink (one tap's jump, ~1e-3..1e-2 of a 121-tap window);
    # RGBA taps carry 7 truncation kinks (3 colour, 4 gradient), gray 5, so
    # about 7/5 as many such pixels
    assert np.median(err) < 1e-4 and np.mean(err < 2e-3) >= (0.93 if colour else 0.95)


def _wave_row_spans(st, P):
    """Width of the other view's texel span of every (64-pixel segment, window
    row) pair, as the image sweep of the adjoint bounds it (planes with a
    finite cost only)."""
    H, W = st.shape[:2]
    nbx = (W + 63) // 64
    xs = np.minimum(np.arange(nbx * 64), W - 1).reshape(nbx, 64)
    act = np.arange(nbx * 64).reshape(nbx, 64) < W
    widths = []
    for y in range(H):
        a, b, c = (st[y, xs, i] for i in range(3))
        for qy in range(max(0, y - P), min(H, y + P + 1)):
            rb = b * np.float32(qy) + c
            xa, xb = a * (xs - P) + rb, a * (xs + P) + rb
            lo, hi = np.minimum(xa, xb), np.maximum(xa, xb)
            ok = act & np.isfinite(lo) & np.isfinite(hi) & (hi >= 0) & (lo < W)
            s0 = np.where(ok, np.maximum(lo, 0), np.inf).min(1)
            s1 = np.where(ok, np.minimum(np.minimum(hi, W).astype(np.int64) + 1, W - 1), -1).max(1)
            widths.append(np.where(s1 >= s0, s1 - np.floor(s0) + 1, 0))
    return np.concatenate(widths)


@pytest.mark.parametrize("colour", [False, True], ids=["gray", "rgba"])
def test_image_gradients_wide_spans_vs_fp64(dev, orc, colour):
    """Random starting planes (disparities up to 320) and a 81x81 window: the
    other view's span of a segment's window row often exceeds the 384 texels
    the kernel sums in LDS (those rows add with global atomics), lanes collide
    on texels and jump back and forth, and this view's row sums need the
    second parking register (2P > 64 columns)."""
    from cvt_amd.synth import make_pair
    pr = make_pair(480, 20, 3, 64.0, 0.01, seed=11, colour=colour)
    P = 40
    L, R = pr.left, pr.right
    GL, GR = orc.gradxy(L), orc.gradxy(R)
    st = orc.init(L, GL, R, GR, P, 320.0, 1)  # disparities drawn from [0, 320]
    spans = _wave_row_spans(st, P)
    assert (spans > 384).mean() > 0.05 and (spans <= 384).mean() > 0.05
    rng = np.random.default_rng(5)
    gout = rng.standard_normal(L.shape[:2]).astype(np.float32)
    _, dI1, dG1, dI2, dG2 = [N(t) for t in dev.cost_backward(T(L), T(GL), T(R), T(GR), T(st),
                                                              T(gout), P)]
    _, rI1, rG1, rI2, rG2 = orc.cost_grad_field_f64(L, GL, R, GR, st, P, gout)
    for got, ref, name in ((dI1, rI1, "dI1"), (dG1, rG1, "dG1"), (dI2, rI2, "dI2"),
                           (dG2, rG2, "dG2")):
        scale = np.abs(ref).max()
        bad = np.abs(got - ref) > 2e-3 * scale + 1e-3 * np.abs(ref)
        print(f"{name}: {bad.mean() * 100:.3f}% outside, scale {scale:.3g}")
        assert bad.mean() < 0.003, (name, bad.mean(), scale)


def test_gradxy_adjoint_dot_product(dev):
    rng = np.random.default_rng(4)
    H, W = 37, 53
    dI = rng.standard_normal((H, W)).astype(np.float32)
    dG = rng.standard_normal((H, W, 4)).astype(np.float32)
    lhs = float(np.sum(N(dev.gradxy(T(dI))).astype(np.float64) * dG))
    rhs = float(np.sum(N(dev.gradxy_backward(T(dG))).astype(np.float64) * dI))
    assert lhs == pytest.approx(rhs, rel=1e-5)


def test_autograd_pair(dev, orc):
    import torch
    from cvt_amd.autograd import WindowCost, window_cost
    from cvt_amd.synth import make_pair
    pr = make_pair(96, 64, 4, 24.0, 0.01, seed=3)
    P = 5
    L, R = T(pr.left), T(pr.right)
    GL, GR = dev.gradxy(L), dev.gradxy(R)
    st = dev.init(L, GL, R, GR, P, 24.0, 1)
    states = st.clone().requires_grad_(True)
    cost = WindowCost.apply(states, L, GL, R, GR, P)
    ok = (cost < 1e5).float()
    (cost * ok).sum().backward()
    ref = dev.cost_backward(L, GL, R, GR, st, ok, P, images=False)[0]
    assert torch.equal(states.grad, ref)
    # through the gradient images to the intensity images
    left = L.clone().requires_grad_(True)
    c2 = window_cost(st, left, R, P)
    (c2 * ok).sum().backward()
    _, dI1, dG1, _, _ = dev.cost_backward(L, GL, R, GR, st, ok, P)
    exp = dev.gradxy_backward(dG1, dI1)
    torch.testing.assert_close(left.grad, exp, rtol=1e-5, atol=1e-6)
    # one gradient step on the planes lowers the summed cost
    with torch.no_grad():
        g = states.grad
        step = torch.zeros_like(st)
        step[..., 2] = -0.05 * torch.sign(g[..., 2])
        c_new = dev.cost(L, GL, R, GR, (st + step).contiguous(), P)
    m = ok.bool() & (c_new < 1e5)
    assert c_new[m].mean() < cost.detach()[m].mean()


def test_autograd_rgba_through_luminance(dev, orc):
    """RGBA inputs: the image gradient reaches R, G, B directly (colour
    terms and weights) and through the gradient image, whose input is the
    luminance 0.2126 R + 0.7152 G + 0.0722 B; alpha gets none."""
    import torch
    from cvt_amd.autograd import window_cost
    from cvt_amd.synth import make_pair
    pr = make_pair(96, 64, 4, 24.0, 0.01, seed=5, colour=True)
    P = 5
    L, R = T(pr.left), T(pr.right)
    GL, GR = dev.gradxy(L), dev.gradxy(R)
    st = dev.init(L, GL, R, GR, P, 24.0, 1)
    left = L.clone().requires_grad_(True)
    c = window_cost(st, left, R, P)
    ok = (c < 1e5).float()
    (c * ok).sum().backward()
    _, dI1, dG1, _, _ = dev.cost_backward(L, GL, R, GR, st, ok, P)
    lum = torch.tensor([0.2126, 0.7152, 0.0722, 0.0], device=L.device)
    exp = dI1 + dev.gradxy_backward(dG1)[..., None] * lum
    assert left.grad.shape == L.shape and not left.grad[..., 3].any()
    torch.testing.assert_close(left.grad, exp, rtol=1e-5, atol=1e-6)


def test_grad_refine_monotone_and_useful(dev, orc, big):
    pr, st = big
    P = 17
    L, R = T(pr.left), T(pr.right)
    GL, GR = dev.gradxy(L), dev.gradxy(R)
    H, W = pr.left.shape
    sm = T(np.zeros((H, W, 4), np.float32))
    s0 = T(st)
    c0 = dev.cost(L, GL, R, GR, s0, P)
    s0[..., 3] = c0
    s1 = dev.grad_refine(s0, L, GL, R, GR, sm, 0.0, P, 192.0, 1, steps=2, radius=0.5)
    c1 = s1[..., 3]
    assert bool((c1 <= c0).all())
    # the stored cost is the cost of the stored plane: the refinement's
    # value-and-gradient sweep sums the same taps in the same order as the
    # forward cost, so the bits agree
    import torch
    assert torch.equal(dev.cost(L, GL, R, GR, s1, P), c1)
    ok = c0 < 1e5
    gain = float((c0[ok] - c1[ok]).mean() / c0[ok].mean())
    moved = float((s1[..., :3] != s0[..., :3]).any(-1).float().mean())
    print(f"grad refine: mean cost -{100 * gain:.2f}%, {100 * moved:.1f}% of planes moved")
    assert gain > 0.005 and moved > 0.2


@pytest.mark.parametrize("lr,theta", [(1, 0.0), (0, 0.0), (1, 0.5)])
def test_grad_refine_matches_oracle(dev, orc, lr, theta):
    """The gradient plane refinement against the oracle's restatement of its
    algorithm (orc_grad_refine: FP64 gradient, float costs), on a field of
    two plain iterations with a smooth prior.  At least 97% of the pixels end
    with the same plane (a, b within 1e-4, x' at the pixel within 1e-3 px).
    The rest differ where the window cost has a kink: near a converged plane
    many taps sit near d = 0, where |d| changes slope, so the float and FP64
    evaluations take different one-sided derivatives (the same effect the
    plane-gradient test allows for 3% of the pixels).  There neither side
    refines better: the oracle's costs of the two sides' planes agree on
    average within 1e-3 relative."""
    import torch
    from cvt_amd.synth import make_pair
    pr = make_pair(200, 120, 4, 24.0, 0.01, seed=31)
    P, dmax = 5, 24.0
    I1, I2 = (pr.left, pr.right) if lr else (pr.right, pr.left)
    res = orc.depthmap(pr.left, pr.right, P, dmax, 2, 1.0, want_states=True, pd_iters=20)
    st = res["left_states"] if lr else res["right_states"]
    H, W = I1.shape
    rng = np.random.default_rng(3)
    sm = np.zeros((H, W, 4), np.float32)
    sm[..., 0:2] = rng.normal(0, 0.1, (H, W, 2))
    sm[..., 2] = rng.uniform(0.2, 0.6, (H, W))
    G1, G2 = orc.gradxy(I1), orc.gradxy(I2)
    so = st.copy()
    so[..., 3] = orc.cost_field(I1, G1, I2, G2, so, P)
    exp = orc.grad_refine(so, I1, G1, I2, G2, sm, theta, P, dmax, lr, 2, 0.5)
    L, R = T(I1), T(I2)
    GL, GR = dev.gradxy(L), dev.gradxy(R)
    sg = T(st)
    sg[..., 3] = dev.cost(L, GL, R, GR, sg, P)
    got = dev.grad_refine(sg, L, GL, R, GR, T(sm), theta, P, dmax, lr, steps=2,
                          radius=0.5).cpu().numpy()
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    xg = got[..., 0] * xs + got[..., 1] * ys + got[..., 2]
    xo = exp[..., 0] * xs + exp[..., 1] * ys + exp[..., 2]
    same = ((np.abs(got[..., 0] - exp[..., 0]) <= 1e-4) &
            (np.abs(got[..., 1] - exp[..., 1]) <= 1e-4) & (np.abs(xg - xo) <= 1e-3))
    moved = float(np.any(exp[..., :3] != so[..., :3], axis=-1).mean())
    print(f"grad refine lr={lr} theta={theta}: {same.mean():.4f} of planes match the oracle; "
          f"{100 * moved:.1f}% moved")
    assert same.mean() >= 0.97 and moved > 0.2
    # both sides' planes costed by the oracle: no side better on average, and
    # the matching planes' costs equal within the window cost's tolerance at
    # the median
    cg = orc.cost_field(I1, G1, I2, G2, got, P)
    ok = exp[..., 3] < 1e5
    assert abs(float((cg[ok] - exp[..., 3][ok]).mean())) <= 1e-3 * float(exp[..., 3][ok].mean())
    c_g, c_o = cg[same & ok], exp[..., 3][same & ok]
    assert np.median(np.abs(c_g - c_o) / np.maximum(np.abs(c_o), 1e-6)) <= 1e-5
