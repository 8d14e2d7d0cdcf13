"""End-to-end parity of PMHuberStereo::depthMap on the 450x375 pair
(8 planes, disparity 0-64, window 21x21, 3 iterations): the HIP path and the
oracle agree within 0.5 px on at least 99% of pixels."""
import numpy as np
import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


def test_depthmap_450x375_matches_oracle(orc):
    import torch
    from cvt_amd import PMHuberStereo, synth
    pr = synth.pair_450x375()
    P, dmax, iters = 10, 64.0, 3
    print(f"depthmap parity: 450x375 P={P} iters={iters}", flush=True)
    gd = PMHuberStereo().depthMap(pr.left, pr.right, P, dmax, iters, 4, 1.0)
    torch.cuda.synchronize()
    gd = gd.cpu().numpy()
    od = orc.depthmap(pr.left, pr.right, P, dmax, iters, 1.0)["dmap"]
    close = np.abs(gd - od) <= 0.5
    frac = close.mean()
    bad_g, bad_o = synth.bad_pixels(gd, pr), synth.bad_pixels(od, pr)
    print(f"within 0.5px: {frac:.4f}; bad-1.0px gpu {bad_g:.2f}% oracle {bad_o:.2f}%")
    assert frac >= 0.99
    assert abs(bad_g - bad_o) < 2.0


def test_depthmap_with_gradient_refinement():
    """The gradient-step refinement inside depthMap runs and does not degrade the map."""
    import torch
    from cvt_amd import _lib, ops, synth
    pr = synth.pair_450x375()
    L, R = torch.from_numpy(pr.left).cuda(), torch.from_numpy(pr.right).cuda()
    base = _lib.default_options(sweep=_lib.SWEEP_REDBLACK, maxdispdiff=1.0)
    grad = _lib.default_options(sweep=_lib.SWEEP_REDBLACK, maxdispdiff=1.0, grad_steps=2)
    d0 = ops.depthmap(L, R, 10, 64.0, 3, 4, 1.0, base)["dmap"].cpu().numpy()
    d1 = ops.depthmap(L, R, 10, 64.0, 3, 4, 1.0, grad)["dmap"].cpu().numpy()
    assert np.isfinite(d1).all()
    b0, b1 = synth.bad_pixels(d0, pr), synth.bad_pixels(d1, pr)
    print(f"bad-1.0px: plain {b0:.3f}%  with 2 gradient steps {b1:.3f}%")
    assert b1 <= b0 + 0.5


def test_depthmap_without_random_neighbourhood_matches_oracle(orc):
    """random_neighbours = 0 (an option of the batched method) against the
    oracle run with the same option: red-black, 200x120, window 11x11."""
    import torch
    from cvt_amd import _lib, ops, synth
    pr = synth.make_pair(200, 120, 5, 24.0, 0.01, seed=31)
    P, dmax, iters = 5, 24.0, 2
    print(f"depthmap parity without random neighbours: 200x120 P={P} iters={iters}", flush=True)
    L, R = torch.from_numpy(pr.left).cuda(), torch.from_numpy(pr.right).cuda()
    opts = _lib.default_options(sweep=_lib.SWEEP_REDBLACK, random_neighbours=0)
    gd = ops.depthmap(L, R, P, dmax, iters, 4, 1.0, opts)["dmap"].cpu().numpy()
    od = orc.depthmap(pr.left, pr.right, P, dmax, iters, 1.0, sweep=1,
                      random_neighbours=0)["dmap"]
    ref = orc.depthmap(pr.left, pr.right, P, dmax, iters, 1.0, sweep=1)["dmap"]
    frac = (np.abs(gd - od) <= 0.5).mean()
    print(f"within 0.5px: {frac:.4f}; differs from the reference setting on "
          f"{(np.abs(od - ref) > 0.5).mean():.4f} of pixels")
    assert frac >= 0.99
    assert (np.abs(od - ref) > 0.5).mean() > 0.001


def test_depthmap_fewer_refinements_matches_oracle(orc):
    """random_refine = 0 and 1 against the oracle run with the same option."""
    import torch
    from cvt_amd import _lib, ops, synth
    pr = synth.make_pair(200, 120, 5, 24.0, 0.01, seed=33)
    P, dmax, iters = 5, 24.0, 2
    L, R = torch.from_numpy(pr.left).cuda(), torch.from_numpy(pr.right).cuda()
    for nref in (0, 1):
        print(f"depthmap parity with random_refine={nref}: 200x120 P={P}", flush=True)
        opts = _lib.default_options(sweep=_lib.SWEEP_REDBLACK, random_refine=nref)
        gd = ops.depthmap(L, R, P, dmax, iters, 4, 1.0, opts)["dmap"].cpu().numpy()
        od = orc.depthmap(pr.left, pr.right, P, dmax, iters, 1.0, sweep=1,
                          random_refine=nref)["dmap"]
        frac = (np.abs(gd - od) <= 0.5).mean()
        print(f"within 0.5px: {frac:.4f}")
        assert frac >= 0.99


@pytest.mark.parametrize("sweep", [0, 1])
def test_depthmap_8_iterations_theta_matches_oracle(orc, sweep):
    """Eight iterations: theta = smoothstep((iter - 5) / 3) becomes 0.259 before
    the last iteration (PMHuberStereo.cpp:264-265), so the smoothness-weighted
    acceptance (pmhstereo.cl:340) and the warm-started Huber-ROF smoothing with
    lambda = 50 theta + 5 take part.  Reference Jacobi sweep and red-black."""
    import torch
    from cvt_amd import _lib, ops, synth
    pr = synth.make_pair(200, 120, 5, 24.0, 0.01, seed=35)
    P, dmax, iters = 5, 24.0, 8
    print(f"depthmap parity, 8 iterations: 200x120 P={P} sweep={sweep}", flush=True)
    L, R = torch.from_numpy(pr.left).cuda(), torch.from_numpy(pr.right).cuda()
    opts = _lib.default_options(sweep=_lib.SWEEP_REDBLACK if sweep else _lib.SWEEP_ALL)
    gd = ops.depthmap(L, R, P, dmax, iters, 4, 1.0, opts)["dmap"].cpu().numpy()
    od = orc.depthmap(pr.left, pr.right, P, dmax, iters, 1.0, sweep=sweep)["dmap"]
    # theta is live: 7 iterations (theta 0 throughout) end elsewhere
    o7 = orc.depthmap(pr.left, pr.right, P, dmax, 7, 1.0, sweep=sweep)["dmap"]
    frac = (np.abs(gd - od) <= 0.5).mean()
    print(f"within 0.5px: {frac:.4f}; bad-1.0px gpu {synth.bad_pixels(gd, pr):.2f}% oracle "
          f"{synth.bad_pixels(od, pr):.2f}%")
    assert frac >= 0.99
    assert not np.array_equal(od, o7)


def test_depthmap_c2_1242x375_matches_oracle(orc):
    """The plain GPU benchmark configuration at its exact settings: 1242x375,
    16 planes, disparity 0-192, noise 0.01, window 35x35, 4 iterations of
    red-black propagation + random refinement, L/R tolerance 1 px; the HIP path
    and the oracle (the same red-black schedule) agree within 0.5 px on at
    least 99% of pixels."""
    import torch
    from cvt_amd import _lib, ops, synth
    pr = synth.make_pair(1242, 375, 16, 192.0, 0.01, seed=2)
    P, dmax, iters = 17, 192.0, 4
    print("depthmap parity C2: 1242x375 P=17 4 iterations red-black, tol 1 px", flush=True)
    L, R = torch.from_numpy(pr.left).cuda(), torch.from_numpy(pr.right).cuda()
    opts = _lib.default_options(sweep=_lib.SWEEP_REDBLACK, maxdispdiff=1.0)
    gd = ops.depthmap(L, R, P, dmax, iters, 4, 1.0, opts)["dmap"].cpu().numpy()
    od = orc.depthmap(pr.left, pr.right, P, dmax, iters, 1.0, maxdispdiff=1.0, sweep=1)["dmap"]
    frac = (np.abs(gd - od) <= 0.5).mean()
    bg, bo = synth.bad_pixels(gd, pr), synth.bad_pixels(od, pr)
    print(f"within 0.5px: {frac:.4f}; bad-1.0px gpu {bg:.3f}% oracle {bo:.3f}%")
    assert frac >= 0.99
    assert abs(bg - bo) < 0.5
