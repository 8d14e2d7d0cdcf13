"""The batched workloads at their full sizes.

64 uneven pairs (W 320-1920, H 240-1080, disparity 32-256, 4-24 planes,
window 35x35, 4 iterations red-black, L/R tolerance 1 px):
  - one pmh_batch_depthmap over all pairs, replayed as a captured graph with
    autotuned shapes, equals pmh_depthmap on each pair alone, bit for bit
    (disparities and both plane fields);
  - the window cost of the largest pair's final plane field matches the
    oracle within 1e-5 relative on 1024 seeded pixels;
  - bad-1.0px over all pairs stays below 1.5% (measured 0.88%).
2880x1988 (32 planes, disparity 0-400, noise 0.02), 8 iterations: theta
becomes 0.259 before the last iteration, so the smoothness term is live:
  - a batch of this pair with a smaller one equals the single-pair run;
  - sampled window costs match the oracle within 1e-5 relative;
  - bad-1.0px below 1.0% (measured 0.46%).
"""
import numpy as np
import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

P = 17


def T(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _plain():
    from cvt_amd import _lib
    return _lib.default_options(sweep=_lib.SWEEP_REDBLACK, maxdispdiff=1.0)


def _sampled_cost_parity(orc, pair, states, n=1024, seed=0):
    """GPU window cost of a plane field against the oracle on n seeded pixels."""
    import torch
    from cvt_amd import ops
    L, R = T(pair.left), T(pair.right)
    GL, GR = ops.gradxy(L), ops.gradxy(R)
    g = ops.cost(L, GL, R, GR, states, P).cpu().numpy()
    st = states.cpu().numpy()
    gl, gr = orc.gradxy(pair.left), orc.gradxy(pair.right)
    np.testing.assert_array_equal(GL.cpu().numpy(), gl)
    H, W = pair.left.shape
    rng = np.random.default_rng(seed)
    ys, xs = rng.integers(0, H, n), rng.integers(0, W, n)
    o = np.array([orc.cost(pair.left, gl, pair.right, gr, int(x), int(y), st[y, x], P)
                  for x, y in zip(xs, ys)], np.float32)
    np.testing.assert_allclose(g[ys, xs], o, rtol=1e-5, atol=1e-7)
    # the field's own stored costs are the same window costs
    np.testing.assert_allclose(st[ys, xs, 3], o, rtol=1e-5, atol=1e-7)
    torch.cuda.synchronize()


def _bad_all(maps, pairs):
    from cvt_amd import synth
    n_bad = n_all = 0
    for m, p in zip(maps, pairs):
        k = int(p.nonocc.sum())
        n_bad += synth.bad_pixels(m, p) * k / 100.0
        n_all += k
    return 100.0 * n_bad / n_all


def test_c4_64_uneven_pairs_batch_equals_single():
    import torch
    from cvt_amd import _lib, ops, synth
    pairs = synth.uneven_batch(64, seed=4)
    dev = [(T(p.left), T(p.right), p.dmax, 1.0) for p in pairs]
    print(f"C4: 64 uneven pairs, {sum(p.left.size for p in pairs) / 1e6:.1f} Mpix", flush=True)
    graph = _lib.default_options(sweep=_lib.SWEEP_REDBLACK, maxdispdiff=1.0, use_graph=1)
    ops.batch_depthmap(dev, P, 4, graph)       # eager, tunes
    batch = ops.batch_depthmap(dev, P, 4, graph, states=True)  # captured and replayed
    g0 = ops.context().graph_stats()
    assert g0[1] >= 1, "the batch was not replayed as a graph"
    maps = []
    for k, (l, r, d, s) in enumerate(dev):
        one = ops.depthmap(l, r, P, d, 4, 4, s, _plain(), states=True)
        for key in ("dmap", "left_states", "right_states"):
            assert torch.equal(one[key], batch[k][key]), (k, key)
        maps.append(batch[k]["dmap"].cpu().numpy())
        if k % 16 == 15:
            print(f"C4: {k + 1}/64 pairs equal", flush=True)
    bad = _bad_all(maps, pairs)
    print(f"C4 bad-1.0px {bad:.3f}%", flush=True)
    assert bad < 1.5
    big = int(np.argmax([p.left.size for p in pairs]))
    import oracle
    _sampled_cost_parity(oracle, pairs[big], batch[big]["left_states"])


def test_c5_2880x1988_8_iterations(orc):
    import torch
    from cvt_amd import ops, synth
    pr = synth.pair_2880x1988(seed=5)
    small = synth.make_pair(640, 480, 8, 96.0, 0.01, seed=9)
    L, R = T(pr.left), T(pr.right)
    print("C5: 2880x1988, 8 iterations", flush=True)
    one = ops.depthmap(L, R, P, 400.0, 8, 4, 1.0, _plain(), states=True)
    batch = ops.batch_depthmap([(L, R, 400.0, 1.0), (T(small.left), T(small.right), 96.0, 1.0)],
                               P, 8, _plain(), states=True)
    for key in ("dmap", "left_states", "right_states"):
        assert torch.equal(one[key], batch[0][key]), key
    bad = synth.bad_pixels(one["dmap"].cpu().numpy(), pr)
    print(f"C5 bad-1.0px {bad:.3f}%", flush=True)
    assert bad < 1.0
    _sampled_cost_parity(orc, pr, one["left_states"], seed=1)
