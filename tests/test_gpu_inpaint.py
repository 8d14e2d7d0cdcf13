"""PMHuberStereo::depthMapInpaint on the HIP path against the oracle.

Bars: occmap, normal_depth, the inpainting Huber-ROF step and its applies
bit-exact; the whole run within 0.5 px on at least 99% of pixels; a batch of
pairs bit-identical to the pairs run alone."""
import numpy as np
import pytest

from test_gpu_kernels import _lr_fields

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
def scene(orc):
    from cvt_amd.synth import make_pair
    pr = make_pair(96, 64, 4, 24.0, 0.01, seed=3)
    L, R = pr.left, pr.right
    return dict(pair=pr, L=L, R=R, GL=orc.gradxy(L), GR=orc.gradxy(R), W=96, H=64, dmax=24.0)


def test_occmap_normal_depth_exact(dev, orc, scene):
    l, r = _lr_fields(orc, scene)
    for lr, (p, q) in ((1, (l, r)), (0, (r, l))):
        for tol in (0.5, 1.0):
            o = orc.occmap(p, q, tol, lr)
            g = N(dev.occmap(T(p), T(q), tol, lr))
            np.testing.assert_array_equal(g, o)
            assert 0.0 < o[..., 0].mean() < 1.0
        np.testing.assert_array_equal(N(dev.normal_depth(T(p), scene["dmax"], lr)),
                                      orc.normal_depth(p, scene["dmax"], lr))


def test_pd_inpaint_exact(dev, orc, scene):
    rng = np.random.default_rng(29)
    H, W = scene["H"], scene["W"]
    f = rng.random((H, W, 4)).astype(np.float32)
    f[..., 3] = 1
    w = orc.weight(scene["L"])
    mask = np.zeros((H, W, 4), np.float32)
    mask[..., :3] = (rng.random((H, W)) < 0.6)[..., None]
    mask[..., 3] = 1
    u0 = rng.random((H, W, 4)).astype(np.float32)
    p0 = (0.8 * rng.standard_normal((H, W, 8))).astype(np.float32)
    p0[..., 3] = p0[..., 7] = 0
    ou, op = orc.pd_step_inpaint(u0, p0, f, w, mask, 7.5)
    gu, gp = dev.pd_step_inpaint(T(u0), T(p0), T(f), T(w), T(mask), 7.5)
    np.testing.assert_array_equal(N(gu), ou)
    np.testing.assert_array_equal(N(gp), op)
    for iters in (1, 2, 25):
        np.testing.assert_array_equal(
            N(dev.pdrof_inpaint(T(u0), T(f), T(w), T(mask), 5.0, iters)),
            orc.pdrof_inpaint(u0, f, w, mask, 5.0, iters))


def test_depthmap_inpaint_matches_oracle(orc):
    from cvt_amd import PMHuberStereo, _lib, synth
    pr = synth.make_pair(200, 120, 6, 32.0, 0.01, seed=23)
    P, dmax, iters = 5, 32.0, 6
    opts = _lib.default_inpaint_options(pd_iters=40)
    gd = N(PMHuberStereo().depthMapInpaint(pr.left, pr.right, P, dmax, iters, 4, opts))
    od = orc.depthmap_inpaint(pr.left, pr.right, P, dmax, iters, pd_iters=40)["dmap"]
    frac = (np.abs(gd - od) <= 0.5).mean()
    print(f"inpaint within 0.5px: {frac:.4f}; bad-1.0px gpu {synth.bad_pixels(gd, pr):.2f}% "
          f"oracle {synth.bad_pixels(od, pr):.2f}%")
    assert frac >= 0.99


def test_inpaint_batch_equals_single(dev):
    import torch
    from cvt_amd import _lib
    from cvt_amd.synth import make_pair
    prs = [make_pair(W, H, 4, d, 0.01, seed=50 + k)
           for k, (W, H, d) in enumerate([(120, 80, 16.0), (97, 61, 12.0)])]
    pairs = [(T(p.left), T(p.right), p.dmax) for p in prs]
    opts = _lib.default_inpaint_options(pd_iters=20, sweep=3, grad_steps=1)
    got = dev.batch_depthmap_inpaint(pairs, 5, 5, opts, states=True)
    for (l, r, d), g in zip(pairs, got):
        ref = dev.depthmap_inpaint(l, r, 5, d, 5, 4, opts, states=True)
        assert torch.equal(ref["dmap"], g["dmap"])
        assert torch.equal(ref["left_states"], g["left_states"])
    with pytest.raises(_lib.PmhError, match="iterations"):
        dev.depthmap_inpaint(pairs[0][0], pairs[0][1], 5, 16.0, 0)


@pytest.mark.parametrize("joint", [0, 1])
def test_inpaint_coarse_batch_equals_single(dev, joint):
    """depthMapInpaint with the coarse-to-fine start (reference view order or
    both views per coarse launch): a batch equals one pair at a time, and the
    Python class's depthMapInpaint passes the option through."""
    import torch
    from cvt_amd import PMHuberStereo, _lib
    from cvt_amd.synth import make_pair
    prs = [make_pair(W, H, 4, d, 0.01, seed=80 + k)
           for k, (W, H, d) in enumerate([(150, 90, 16.0), (101, 77, 12.0)])]
    pairs = [(T(p.left), T(p.right), p.dmax) for p in prs]
    kw = dict(pd_iters=20, sweep=3, grad_steps=1, coarse_levels=1, coarse_iters=2,
              coarse_pd_iters=10, coarse_joint_views=joint)
    opts = _lib.default_inpaint_options(**kw)
    got = dev.batch_depthmap_inpaint(pairs, 5, 2, opts, states=True)
    for (l, r, d), g in zip(pairs, got):
        ref = dev.depthmap_inpaint(l, r, 5, d, 2, 4, opts, states=True)
        assert torch.equal(ref["dmap"], g["dmap"])
        assert torch.equal(ref["left_states"], g["left_states"])
    # the class keeps the reference's 100 smoother steps and takes the rest
    stereo = PMHuberStereo(_lib.default_options(**kw))
    fac = stereo.depthMapInpaint(pairs[0][0], pairs[0][1], 5, pairs[0][2], 2)
    ref = dev.depthmap_inpaint(pairs[0][0], pairs[0][1], 5, pairs[0][2], 2, 4,
                               _lib.default_inpaint_options(**dict(kw, pd_iters=100)))
    assert torch.equal(fac, ref["dmap"])
