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
