"""Parity of each HIP kernel (through the C-ABI) with the oracle on seeded inputs.

Bars: bit-exact for integer work (RNG draws, view-propagation slots) and for
kernels whose arithmetic is fully specified (gradients, weights, plane
algebra, consistency, fill, Huber-ROF); the window cost within 1e-5
relative (the kernel's exp is the hardware exp, the oracle's is libm);
propagation choosing identical planes except near cost ties (1e-6).
"""
import numpy as np
import pytest

from proputil import compare_choice as _compare_choice

pytestmark = pytest.mark.gpu

RTOL_COST = 1e-5
TIE = 1e-6


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU test needs a HIP device"
    from cvt_amd import ops
    return ops


def T(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def N(t):
    return t.detach().cpu().numpy()


@pytest.fixture(scope="module")
def scene(orc):
    from cvt_amd.synth import make_pair
    pr = make_pair(96, 64, 4, 24.0, 0.01, seed=3)
    L, R = pr.left, pr.right
    return dict(pair=pr, L=L, R=R, GL=orc.gradxy(L), GR=orc.gradxy(R), W=96, H=64, dmax=24.0)


def test_rng_draws_bitexact(dev, orc):
    W, H = 40, 12
    for it, stride, count in [(0, 3, 3), (2, 45, 21), (7, 45, 21)]:
        g = N(dev.rng_draws(W, H, it, stride, count))
        for y in range(H):
            for x in range(W):
                ref = orc.rng_draws(y * W + x + it, stride, count)
                assert np.array_equal(g[y, x], ref), (x, y, it, stride)


def test_gradxy_weight_bitexact(dev, orc, scene):
    for img in (scene["L"], scene["R"]):
        np.testing.assert_array_equal(N(dev.gradxy(T(img))), orc.gradxy(img))
        np.testing.assert_array_equal(N(dev.weight(T(img))), orc.weight(img))


def _random_planes(rng, W, H, dmax, lr=1):
    st = np.zeros((H, W, 4), np.float32)
    st[..., 0] = 1 + rng.uniform(-0.3, 0.3, (H, W))
    st[..., 1] = rng.uniform(-0.2, 0.2, (H, W))
    xs = np.arange(W)[None, :]
    st[..., 2] = -(st[..., 0] - 1) * xs - rng.uniform(0, dmax, (H, W))
    return st


@pytest.mark.parametrize("P", [0, 3, 5, 17])
def test_cost_field(dev, orc, scene, P):
    rng = np.random.default_rng(P)
    W, H = scene["W"], scene["H"]
    st = _random_planes(rng, W, H, scene["dmax"])
    st[0, 0, 0] = np.nan
    st[1, 1, 2] = np.inf
    args = (scene["L"], scene["GL"], scene["R"], scene["GR"])
    g = N(dev.cost(*[T(a) for a in args], T(st), P))
    o = orc.cost_field(*args, st, P)
    np.testing.assert_allclose(g, o, rtol=RTOL_COST, atol=1e-7)
    assert g[0, 0] == 1e5 and g[1, 1] == 1e5


@pytest.mark.parametrize("lr", [1, 0])
def test_init(dev, orc, scene, lr):
    P = 5
    a = (scene["L"], scene["GL"], scene["R"], scene["GR"]) if lr else \
        (scene["R"], scene["GR"], scene["L"], scene["GL"])
    g = N(dev.init(*[T(x) for x in a], P, scene["dmax"], lr))
    o = orc.init(*a, P, scene["dmax"], lr)
    np.testing.assert_array_equal(g[..., :3], o[..., :3])
    np.testing.assert_allclose(g[..., 3], o[..., 3], rtol=RTOL_COST, atol=1e-7)


def _prop_setup(orc, scene, lr, seed):
    rng = np.random.default_rng(seed)
    P, W, H, dmax = 5, scene["W"], scene["H"], scene["dmax"]
    a = (scene["L"], scene["GL"], scene["R"], scene["GR"]) if lr else \
        (scene["R"], scene["GR"], scene["L"], scene["GL"])
    old = orc.init(*a, P, dmax, lr)
    # smooth prior: a noisy normal/depth field as the Huber-ROF output would be
    sm = np.zeros((H, W, 4), np.float32)
    sm[..., 0:2] = rng.uniform(-0.2, 0.2, (H, W, 2))
    sm[..., 2] = rng.uniform(0, 1, (H, W))
    sm[..., 3] = 1
    other = orc.init(*(a[2], a[3], a[0], a[1]), P, dmax, 1 - lr)
    vb = orc.viewbuf_build(other)
    draws = rng.random((H, W, 21)).astype(np.float32)
    return a, old, sm, other, vb, draws


def _draws_of(orc, W, H, it):
    return np.stack([np.stack([orc.rng_draws(y * W + x + it, 45, 21) for x in range(W)])
                     for y in range(H)])


@pytest.mark.parametrize("lr", [1, 0])
@pytest.mark.parametrize("theta", [0.0, 0.5])
def test_propagate_jacobi_supplied_draws(dev, orc, scene, lr, theta):
    a, old, sm, other, vb, draws = _prop_setup(orc, scene, lr, 11 + lr)
    P, dmax = 5, scene["dmax"]
    o = orc.propagate(old, *a, sm, theta, P, dmax, lr, 2, viewin=vb, viewsrc=other, draws=draws)
    g = N(dev.propagate(T(old), *[T(x) for x in a], T(sm), theta, P, dmax, lr, 2,
                        viewin=T(vb), viewsrc=T(other), draws=T(draws)))
    ctx = dict(old=old, out=old, sm=sm, theta=theta, dmax=dmax, draws=draws, vb=vb, vsrc=other,
               sweep=0)
    _compare_choice(orc, a, g, o, P, lr, f"jacobi lr={lr} theta={theta}", ctx)


@pytest.mark.parametrize("lr", [1, 0])
def test_propagate_own_rng_red_black(dev, orc, scene, lr):
    """Red pass, then black pass, each checked on identical inputs (the black
    pass reads the red pixels the GPU wrote)."""
    a, old, sm, other, vb, _ = _prop_setup(orc, scene, lr, 21 + lr)
    P, dmax, it = 5, scene["dmax"], 3
    draws = _draws_of(orc, scene["W"], scene["H"], it)
    gout = T(old.copy())
    dev.propagate(T(old), *[T(x) for x in a], T(sm), 0.0, P, dmax, lr, it, viewin=T(vb),
                  viewsrc=T(other), sweep=1, out=gout)
    g_red = N(gout)
    o_red = orc.propagate(old, *a, sm, 0.0, P, dmax, lr, it, viewin=vb, viewsrc=other,
                          sweep=1, out=old.copy())
    ctx = dict(old=old, out=old, sm=sm, theta=0.0, dmax=dmax, draws=draws, vb=vb, vsrc=other,
               sweep=1)
    _compare_choice(orc, a, g_red, o_red, P, lr, f"red lr={lr}", ctx)
    dev.propagate(T(old), *[T(x) for x in a], T(sm), 0.0, P, dmax, lr, it, viewin=T(vb),
                  viewsrc=T(other), sweep=2, out=gout)
    g_blk = N(gout)
    o_blk = orc.propagate(old, *a, sm, 0.0, P, dmax, lr, it, viewin=vb, viewsrc=other,
                          sweep=2, out=g_red.copy())
    ctx.update(out=g_red, sweep=2)
    _compare_choice(orc, a, g_blk, o_blk, P, lr, f"black lr={lr}", ctx)


def test_viewbuf_exact(dev, orc, scene):
    rng = np.random.default_rng(5)
    st = _random_planes(rng, scene["W"], scene["H"], scene["dmax"])
    st[3, 3, :3] = np.nan
    g = N(dev.viewbuf_build(T(st)))
    o = orc.viewbuf_build(st)
    np.testing.assert_array_equal(g[..., 0], o[..., 0])
    n = np.minimum(o[..., 0], 4)
    for k in range(4):
        m = n > k
        np.testing.assert_array_equal(g[..., 1 + k][m], o[..., 1 + k][m])


def _lr_fields(orc, scene):
    a = (scene["L"], scene["GL"], scene["R"], scene["GR"])
    b = (scene["R"], scene["GR"], scene["L"], scene["GL"])
    sm = np.zeros((scene["H"], scene["W"], 4), np.float32)
    l = orc.init(*a, 5, scene["dmax"], 1)
    r = orc.init(*b, 5, scene["dmax"], 0)
    for it in range(2):
        l = orc.propagate(l, *a, sm, 0.0, 5, scene["dmax"], 1, it)
        r = orc.propagate(r, *b, sm, 0.0, 5, scene["dmax"], 0, it)
    return l, r


def test_consistency_fill_exact(dev, orc, scene):
    l, r = _lr_fields(orc, scene)
    dmax = scene["dmax"]
    for lr, (p, q) in ((1, (l, r)), (0, (r, l))):
        for tol in (0.5, 1.0):
            oc = orc.consistency(p, q, tol, 10.0, lr)
            gc = N(dev.consistency(T(p), T(q), tol, 10.0, lr))
            np.testing.assert_array_equal(gc, oc)
            assert 0.0 < (oc[..., 0] == 0).mean() < 1.0  # some pixels rejected, not all
            of = orc.fill_state(oc, dmax, lr)
            gf = N(dev.fill_state(T(oc), dmax, lr))
            np.testing.assert_array_equal(gf, of)
    oc = orc.consistency(l, r, 0.5, 10.0, 1)
    np.testing.assert_array_equal(N(dev.fill_depthmap(T(oc), 1.0 / dmax)),
                                  orc.fill_depthmap(oc, 1.0 / dmax))
    np.testing.assert_array_equal(N(dev.fill_normalmap(T(oc))), orc.fill_normalmap(oc))


def test_pd_exact(dev, orc, scene):
    rng = np.random.default_rng(9)
    H, W = scene["H"], scene["W"]
    f = rng.random((H, W, 4)).astype(np.float32)
    f[..., 3] = 1
    w = orc.weight(scene["L"])
    u0 = rng.random((H, W, 4)).astype(np.float32)
    p0 = (0.5 * rng.standard_normal((H, W, 8))).astype(np.float32)
    p0[..., 3] = p0[..., 7] = 0
    ou, op = orc.pd_step(u0, p0, f, w, 7.5)
    gu, gp = dev.pd_step(T(u0), T(p0), T(f), T(w), 7.5)
    np.testing.assert_array_equal(N(gu), ou)
    np.testing.assert_array_equal(N(gp), op)
    for iters in (1, 2, 25):
        np.testing.assert_array_equal(N(dev.pdrof(T(u0), T(f), T(w), 5.0, iters)),
                                      orc.pdrof(u0, f, w, 5.0, iters))
