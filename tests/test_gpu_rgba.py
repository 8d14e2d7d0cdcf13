"""RGBA inputs on the HIP path (PMH_FORMAT_RGBA, through the C-ABI).

Bars as for gray inputs: gradients, weights and init planes bit-exact with the
oracle, window cost within 1e-5, propagation choosing identical planes except
near cost ties (1e-6), depthMap within 0.5 px on at least 99% of pixels.  An
RGBA pair whose R = G = B are the gray images gives the gray path's gradients
bit for bit, its window costs to rounding (the support-weight exponent is
folded per format: |d| (-3 F log2 e) for gray, (3 |d|) (-F log2 e) for RGBA)
and its depth maps within 0.5 px."""
import numpy as np
import pytest

from test_gpu_kernels import RTOL_COST, _compare_choice, _draws_of, _random_planes

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


@pytest.fixture(scope="module")
def cscene(orc):
    from cvt_amd.synth import make_pair
    pr = make_pair(96, 64, 4, 24.0, 0.01, seed=13, colour=True)
    L, R = pr.left, pr.right
    return dict(pair=pr, L=L, R=R, GL=orc.gradxy(L), GR=orc.gradxy(R), W=96, H=64, dmax=24.0)


def test_rgba_gradxy_weight_pyr_bitexact(dev, orc, cscene):
    for img in (cscene["L"], cscene["R"]):
        np.testing.assert_array_equal(N(dev.gradxy(T(img))), orc.gradxy(img))
        np.testing.assert_array_equal(N(dev.weight(T(img))), orc.weight(img))
    img = cscene["L"][:63, :95]
    got = N(dev.pyr_down(T(img)))
    H, W = img.shape[:2]
    y1, x1 = np.minimum(2 * np.arange((H + 1) // 2) + 1, H - 1), \
        np.minimum(2 * np.arange((W + 1) // 2) + 1, W - 1)
    y0, x0 = 2 * np.arange((H + 1) // 2), 2 * np.arange((W + 1) // 2)
    exp = np.float32(0.25) * ((img[y0][:, x0] + img[y0][:, x1]) + (img[y1][:, x0] + img[y1][:, x1]))
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("P", [3, 17])
def test_rgba_cost_field(dev, orc, cscene, P):
    rng = np.random.default_rng(100 + P)
    st = _random_planes(rng, cscene["W"], cscene["H"], cscene["dmax"])
    args = (cscene["L"], cscene["GL"], cscene["R"], cscene["GR"])
    g = N(dev.cost(*[T(a) for a in args], T(st), P))
    o = orc.cost_field(*args, st, P)
    np.testing.assert_allclose(g, o, rtol=RTOL_COST, atol=1e-7)


@pytest.mark.parametrize("lr", [1, 0])
def test_rgba_init(dev, orc, cscene, lr):
    P = 5
    a = (cscene["L"], cscene["GL"], cscene["R"], cscene["GR"]) if lr else \
        (cscene["R"], cscene["GR"], cscene["L"], cscene["GL"])
    g = N(dev.init(*[T(x) for x in a], P, cscene["dmax"], lr))
    o = orc.init(*a, P, cscene["dmax"], lr)
    np.testing.assert_array_equal(g[..., :3], o[..., :3])
    np.testing.assert_allclose(g[..., 3], o[..., 3], rtol=RTOL_COST, atol=1e-7)


@pytest.mark.parametrize("lr", [1, 0])
def test_rgba_propagate_red_black(dev, orc, cscene, lr):
    rng = np.random.default_rng(31 + lr)
    P, W, H, dmax, it = 5, cscene["W"], cscene["H"], cscene["dmax"], 1
    a = (cscene["L"], cscene["GL"], cscene["R"], cscene["GR"]) if lr else \
        (cscene["R"], cscene["GR"], cscene["L"], cscene["GL"])
    old = orc.init(*a, P, dmax, lr)
    sm = np.zeros((H, W, 4), np.float32)
    sm[..., 0:2] = rng.uniform(-0.2, 0.2, (H, W, 2))
    sm[..., 2] = rng.uniform(0, 1, (H, W))
    other = orc.init(a[2], a[3], a[0], a[1], P, dmax, 1 - lr)
    vb = orc.viewbuf_build(other)
    draws = _draws_of(orc, W, H, it)
    gout = T(old.copy())
    dev.propagate(T(old), *[T(x) for x in a], T(sm), 0.0, P, dmax, lr, it, viewin=T(vb),
                  viewsrc=T(other), sweep=1, out=gout)
    g_red = N(gout)
    o_red = orc.propagate(old, *a, sm, 0.0, P, dmax, lr, it, viewin=vb, viewsrc=other, sweep=1,
                          out=old.copy())
    ctx = dict(old=old, out=old, sm=sm, theta=0.0, dmax=dmax, draws=draws, vb=vb, vsrc=other,
               sweep=1)
    _compare_choice(orc, a, g_red, o_red, P, lr, f"rgba red lr={lr}", ctx)


def test_gray_replicated_rgba_matches_gray(dev):
    import torch
    from cvt_amd import _lib
    from cvt_amd.synth import make_pair
    pr = make_pair(130, 77, 4, 20.0, 0.01, seed=17)

    def rgba(g):
        return np.stack([g, g, g, np.ones_like(g)], -1)
    L, R = T(pr.left), T(pr.right)
    L4, R4 = T(rgba(pr.left)), T(rgba(pr.right))
    G, GR = dev.gradxy(L), dev.gradxy(R)
    assert torch.equal(dev.gradxy(L4), G) and torch.equal(dev.gradxy(R4), GR)
    opts = _lib.default_options(sweep=3, pd_iters=20, coarse_levels=1, coarse_iters=2)
    a = dev.depthmap(L, R, 5, 20.0, 2, 4, 1.0, opts, normals=True, states=True)
    b = dev.depthmap(L4, R4, 5, 20.0, 2, 4, 1.0, opts, normals=True, states=True)
    # window costs of the same planes
    st = a["left_states"]
    c1 = N(dev.cost(L, G, R, GR, st, 5))
    c4 = N(dev.cost(L4, G, R4, GR, st, 5))
    np.testing.assert_allclose(c4, c1, rtol=RTOL_COST, atol=1e-7)
    frac = (np.abs(N(a["dmap"]) - N(b["dmap"])) <= 0.5).mean()
    print(f"gray vs replicated RGBA depthMap within 0.5px: {frac:.4f}")
    assert frac >= 0.99


def test_rgba_depthmap_matches_oracle(orc):
    import torch
    from cvt_amd import PMHuberStereo, synth
    pr = synth.make_pair(200, 120, 6, 32.0, 0.01, seed=19, colour=True)
    P, dmax, iters = 5, 32.0, 3
    gd = N(PMHuberStereo().depthMap(pr.left, pr.right, P, dmax, iters, 4, 1.0))
    od = orc.depthmap(pr.left, pr.right, P, dmax, iters, 1.0)["dmap"]
    frac = (np.abs(gd - od) <= 0.5).mean()
    print(f"rgba depthmap within 0.5px: {frac:.4f}; bad-1.0px gpu "
          f"{synth.bad_pixels(gd, pr):.2f}% oracle {synth.bad_pixels(od, pr):.2f}%")
    assert frac >= 0.99


def test_rgba_grad_refine_stored_cost(dev, orc, cscene):
    import torch
    from cvt_amd import _lib
    P = 5
    L, R, GL, GR = (T(cscene[k]) for k in ("L", "R", "GL", "GR"))
    st = T(orc.init(cscene["L"], cscene["GL"], cscene["R"], cscene["GR"], P, 24.0, 1))
    st[..., 3] = dev.cost(L, GL, R, GR, st, P)   # stored costs from this path
    sm = torch.zeros_like(st)
    s1 = dev.grad_refine(st, L, GL, R, GR, sm, 0.0, P, 24.0, 1, steps=2, radius=0.5)
    assert bool((s1[..., 3] <= st[..., 3]).all())
    assert torch.equal(dev.cost(L, GL, R, GR, s1, P), s1[..., 3])
