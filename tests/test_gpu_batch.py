"""Batched depthMap (pmh_batch_depthmap), graph replay and the autotuner.

A batch shares every launch between its pairs, but each pair's arithmetic is
the single-pair path's: the results must be bit-identical to running
pmh_depthmap on each pair alone.  The launch shapes the autotuner picks (LDS
stage width, pixels per wavefront) and graph replay must not change a bit
either.
"""
import os

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


SHAPES = [(96, 64, 3, 16.0), (161, 90, 4, 24.0), (130, 77, 5, 32.0), (64, 40, 2, 12.0)]


@pytest.fixture(scope="module")
def pairs():
    from cvt_amd.synth import make_pair
    out = []
    for k, (W, H, n, dmax) in enumerate(SHAPES):
        pr = make_pair(W, H, n, dmax, 0.01, seed=20 + k)
        out.append((T(pr.left), T(pr.right), dmax))
    return out


def _single(dev, pairs, P, it, opts):
    return [dev.depthmap(l, r, P, d, it, 4, -1.0, opts, normals=True, states=True)
            for l, r, d in pairs]


def _same(a, b):
    import torch
    for k in ("dmap", "normals", "left_states", "right_states"):
        assert torch.equal(a[k], b[k]), k


@pytest.mark.parametrize("sweep", [0, 3])
def test_batch_equals_single_pairs(dev, pairs, sweep):
    from cvt_amd import _lib
    opts = _lib.default_options(sweep=sweep, pd_iters=20, grad_steps=1 if sweep else 0)
    ref = _single(dev, pairs, 5, 3, opts)
    got = dev.batch_depthmap(pairs, 5, 3, opts, normals=True, states=True)
    for a, b in zip(got, ref):
        _same(a, b)


def test_graph_replays_match_eager(dev, pairs):
    import torch
    from cvt_amd import _lib
    from cvt_amd.synth import make_pair
    eager = _lib.default_options(sweep=3, pd_iters=30, grad_steps=1)
    graph = _lib.default_options(sweep=3, pd_iters=30, grad_steps=1, use_graph=1)
    ref = dev.batch_depthmap(pairs, 5, 2, eager, normals=True, states=True)
    ctx = dev.context()
    g0, n0 = ctx.graph_stats()
    for rep in range(3):
        got = dev.batch_depthmap(pairs, 5, 2, graph, normals=True, states=True)
        for a, b in zip(got, ref):
            _same(a, b)
        # unrelated eager work and a synchronization between replays
        x = torch.randn(1 << 20, device="cuda")
        assert torch.isfinite((x * 2).sum())
        torch.cuda.synchronize()
    g1, n1 = ctx.graph_stats()
    assert n1 - n0 >= 2, "the graph was not replayed"
    assert g1 - g0 <= 1, "one launch list needs one graph"
    # new image contents with the same shapes replay the same graph
    new = []
    for k, (W, H, n, dmax) in enumerate(SHAPES):
        pr = make_pair(W, H, n, dmax, 0.01, seed=40 + k)
        new.append((T(pr.left), T(pr.right), dmax))
    ref2 = dev.batch_depthmap(new, 5, 2, eager, normals=True, states=True)
    got2 = dev.batch_depthmap(new, 5, 2, graph, normals=True, states=True)
    for a, b in zip(got2, ref2):
        _same(a, b)
    g2, n2 = ctx.graph_stats()
    assert g2 == g1 and n2 == n1 + 1


def _fresh_run(dev, pairs, opts, autotune, sample=0, tuning=None):
    """A batch run on a context of its own (contexts are per stream), so no
    shape tuned by an earlier test applies; returns the results and the
    context's tuning lines."""
    import tempfile

    import torch
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        ctx = dev.context()
        ctx.autotune(autotune)
        ctx.autotune_sample(sample)
        if tuning:
            ctx.tuning_load(tuning)
        res = dev.batch_depthmap(pairs, 5, 2, opts, states=True)
        torch.cuda.synchronize()
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "t.txt")
            ctx.tuning_save(path)
            lines = [ln for ln in open(path).read().split("\n") if ln.strip()]
    return res, lines


def test_autotuner_choices_do_not_change_results(dev, pairs, tmp_path):
    """A run with the default shapes (autotuning off, nothing loaded) against a
    run that tunes every shape class during the call, each on a fresh context."""
    import torch
    from cvt_amd import _lib
    opts = _lib.default_options(sweep=3, pd_iters=10)
    ref, none = _fresh_run(dev, pairs, opts, autotune=False)
    assert none == []  # nothing was tuned: the defaults ran
    got, lines = _fresh_run(dev, pairs, opts, autotune=True)
    for a, b in zip(got, ref):
        assert torch.equal(a["dmap"], b["dmap"])
        assert torch.equal(a["left_states"], b["left_states"])
        assert torch.equal(a["right_states"], b["right_states"])
    keys = [ln.split()[0] for ln in lines]
    assert any(k.startswith("propagate|p5|c1|s1|") and k.endswith("|r2|n6") for k in keys), keys
    assert any(k.startswith("pd_steps|") for k in keys), keys
    for ln in lines:
        _, st, ch = ln.split()
        assert int(st) in (0, 1, 2, 3, 4, 5, 512, 704, 896) and int(ch) in (1, 2, 4, 8)
    path = str(tmp_path / "tuning.txt")
    open(path, "w").write("\n".join(lines) + "\n")
    fresh = _lib.Context(0)
    fresh.tuning_load(path)
    path2 = str(tmp_path / "tuning2.txt")
    fresh.tuning_save(path2)
    assert sorted(ln for ln in open(path2).read().split("\n") if ln.strip()) == sorted(lines)
    fresh.close()
    # the loaded choices, tuning off, give the same bits again
    again, _ = _fresh_run(dev, pairs, opts, autotune=False, tuning=path)
    for a, b in zip(again, ref):
        assert torch.equal(a["left_states"], b["left_states"])


def test_autotuner_marginal_scoring(dev, pairs):
    """A sample of 8 tiles makes every launch of the batch at least twice the
    sample, so every candidate is scored by its marginal time (the path the
    big batched workloads take); the tuned run equals the default one."""
    import torch
    from cvt_amd import _lib
    opts = _lib.default_options(sweep=3, pd_iters=10)
    ref, _ = _fresh_run(dev, pairs, opts, autotune=False)
    got, lines = _fresh_run(dev, pairs, opts, autotune=True, sample=8)
    for a, b in zip(got, ref):
        assert torch.equal(a["left_states"], b["left_states"])
        assert torch.equal(a["dmap"], b["dmap"])
    assert any(ln.startswith("propagate|") for ln in lines)
    ctx = _lib.Context(0)
    for bad in (1, 7, -3):
        with pytest.raises(_lib.PmhError):
            ctx.autotune_sample(bad)
    ctx.autotune_sample(0)
    ctx.close()


def _forced_stages(dev, pairs, P, it, opts, tmp_path, stages):
    """Results of the same batch with every propagate launch forced to each
    stage code."""
    ctx = dev.context()
    dev.batch_depthmap(pairs, P, it, opts)  # tunes every shape class of this run
    saved = str(tmp_path / "tuned.txt")
    ctx.tuning_save(saved)
    lines = [ln.split() for ln in open(saved).read().split("\n") if ln.strip()]
    assert any(k.startswith("propagate|") for k, _, _ in lines)
    got = {}
    try:
        ctx.autotune(False)
        for stage in stages:
            forced = tmp_path / f"forced{stage}.txt"
            forced.write_text("".join(
                f"{k} {stage if k.startswith('propagate|') else s} {c}\n" for k, s, c in lines))
            ctx.tuning_load(str(forced))
            got[stage] = dev.batch_depthmap(pairs, P, it, opts, normals=True, states=True)
    finally:
        ctx.tuning_load(saved)
        ctx.autotune(True)
    return got


@pytest.mark.parametrize("stage", ["4", "2", "896"])
def test_persistent_blocks_match(dev, pairs, tmp_path, stage):
    """Every tunable launch (propagate, fused Huber-ROF steps) with persistent
    blocks taking tiles from per-XCD queues (chunk 0) gives the bits of one
    tile per block, eagerly and in three replays of the captured graph (the
    last block of each launch clears the queue counters for the next)."""
    from cvt_amd import _lib
    ctx = dev.context()
    opts = _lib.default_options(sweep=3, pd_iters=10, grad_steps=1)
    dev.batch_depthmap(pairs, 5, 2, opts)  # tunes every shape class of this run
    saved = str(tmp_path / "tuned.txt")
    ctx.tuning_save(saved)
    lines = [ln.split() for ln in open(saved).read().split("\n") if ln.strip()]
    assert any(k.startswith("pd_steps|") for k, _, _ in lines)

    def forced(chunk):
        path = tmp_path / f"forced{stage}_{chunk}.txt"
        path.write_text("".join(
            f"{k} {stage if k.startswith('propagate|') else s} {chunk}\n" for k, s, _ in lines))
        return str(path)

    try:
        ctx.autotune(False)
        ctx.tuning_load(forced(1))
        ref = dev.batch_depthmap(pairs, 5, 2, opts, normals=True, states=True)
        ctx.tuning_load(forced(0))
        for use_graph in (0, 1, 1, 1):
            o = _lib.default_options(sweep=3, pd_iters=10, grad_steps=1, use_graph=use_graph)
            got = dev.batch_depthmap(pairs, 5, 2, o, normals=True, states=True)
            for a, b in zip(got, ref):
                _same(a, b)
    finally:
        ctx.tuning_load(saved)
        ctx.autotune(True)


@pytest.mark.parametrize("sweep", [0, 3])
def test_split_candidates_kernel_matches_sequential(dev, pairs, tmp_path, sweep):
    """Stage code 1 runs each row segment's candidates on four waves (costs in
    parallel, folds in order), stage code 2 also costs each wave's candidates
    in one sweep of the window, code 3 is code 2 at a narrower stage: the same
    planes bit for bit as one wave."""
    from cvt_amd import _lib
    opts = _lib.default_options(sweep=sweep, pd_iters=10, grad_steps=1 if sweep else 0)
    got = _forced_stages(dev, pairs, 5, 2, opts, tmp_path, ("1", "2", "3", "4", "5", "896"))
    for stage in ("1", "2", "3", "4", "5"):
        for a, b in zip(got[stage], got["896"]):
            _same(a, b)


@pytest.mark.parametrize("fmt", ["gray", "rgba"])
def test_multi_plane_cost_on_wide_rows(dev, tmp_path, fmt):
    """A 1100-px-wide pair with a 35x35 window and disparities up to 300: the
    union of a wave's planes often spans more than the LDS stage (then each
    plane takes the row alone), and the right view's random planes are steep.
    Every path of the multi-plane cost gives the one-wave kernel's bits."""
    from cvt_amd import _lib
    from cvt_amd.synth import make_pair
    pr = make_pair(1100, 56, 6, 300.0, 0.01, seed=41)
    l, r = pr.left, pr.right
    if fmt == "rgba":
        l = np.stack([l, 0.8 * l + 0.1, l * l, np.ones_like(l)], -1)
        r = np.stack([r, 0.8 * r + 0.1, r * r, np.ones_like(r)], -1)
    wide = [(T(l.astype(np.float32)), T(r.astype(np.float32)), 300.0)]
    opts = _lib.default_options(sweep=3, pd_iters=10)
    one_wave = "896" if fmt == "gray" else "512"
    got = _forced_stages(dev, wide, 17, 2, opts, tmp_path, ("2", "1", "3", "4", "5", one_wave))
    ref = got[one_wave]
    for stage in ("1", "2", "3", "4", "5"):  # 4 and 5 are gray only (RGBA: the default)
        _same(got[stage][0], ref[0])


def test_coarse_random_refine(dev, pairs):
    """coarse_random_refine sets the coarse levels' randomized refinements:
    -1 (the default) is random_refine's value, and the batched method's setting
    (none at full resolution, the reference's 2 at the coarse level) gives each
    pair the bits of its own run."""
    from cvt_amd import _lib
    base = dict(sweep=3, pd_iters=10, coarse_levels=1, coarse_iters=2, coarse_pd_iters=5,
                grad_steps=1)
    same = dev.batch_depthmap(pairs, 5, 1, _lib.default_options(**base, random_refine=1),
                              normals=True, states=True)
    expl = dev.batch_depthmap(pairs, 5, 1, _lib.default_options(
        **base, random_refine=1, coarse_random_refine=1), normals=True, states=True)
    for a, b in zip(same, expl):
        _same(a, b)
    opts = _lib.default_options(**base, random_refine=0, coarse_random_refine=2)
    got = dev.batch_depthmap(pairs, 5, 1, opts, normals=True, states=True)
    for a, b in zip(got, _single(dev, pairs, 5, 1, opts)):
        _same(a, b)
    import torch
    assert any(not torch.equal(a["left_states"], b["left_states"]) for a, b in zip(got, same))
    with pytest.raises(_lib.PmhError, match="coarse_random_refine"):
        dev.batch_depthmap(pairs, 5, 1, _lib.default_options(coarse_random_refine=3))


@pytest.mark.parametrize("sweep", [0, 3])
def test_random_neighbours_off(dev, pairs, tmp_path, sweep):
    """random_neighbours = 0 leaves the six random-neighbourhood candidates out
    of every pass: each kernel shape (the one-wave kernel, the four-wave kernel
    with one sweep per candidate, the shared-sweep kernels) gives the same bits,
    a batch gives each pair its own run's bits, and the result differs from the
    reference setting."""
    import torch
    from cvt_amd import _lib
    opts = _lib.default_options(sweep=sweep, pd_iters=10, random_neighbours=0, grad_steps=1)
    got = _forced_stages(dev, pairs, 5, 2, opts, tmp_path, ("1", "2", "3", "4", "5", "896"))
    for stage in ("1", "2", "3", "4", "5"):
        for a, b in zip(got[stage], got["896"]):
            _same(a, b)
    for a, b in zip(dev.batch_depthmap(pairs, 5, 2, opts, normals=True, states=True),
                    _single(dev, pairs, 5, 2, opts)):
        _same(a, b)
    ref = dev.batch_depthmap(pairs, 5, 2, _lib.default_options(sweep=sweep, pd_iters=10,
                                                               grad_steps=1), states=True)
    assert any(not torch.equal(a["left_states"], b["left_states"])
               for a, b in zip(got["896"], ref))
    with pytest.raises(_lib.PmhError, match="random_neighbours"):
        dev.batch_depthmap(pairs, 5, 1, _lib.default_options(random_neighbours=3))
    with pytest.raises(_lib.PmhError, match="coarse_random_neighbours"):
        dev.batch_depthmap(pairs, 5, 1, _lib.default_options(coarse_random_neighbours=2))


@pytest.mark.parametrize("W,H", [(40, 1), (3, 2)])
def test_tiny_images_tune_and_run(dev, W, H):
    """Launches of a single tile (one row segment) go through the autotuner
    like any other (its half-size sample is only taken for big launches)."""
    from cvt_amd import _lib
    rng = np.random.default_rng(10 * W + H)
    left = rng.random((H, W)).astype(np.float32)
    right = np.roll(left, -1, axis=1)
    res = dev.batch_depthmap([(T(left), T(right), 4.0)], 2, 1,
                             _lib.default_options(sweep=3, pd_iters=5), states=True)
    assert tuple(res[0]["dmap"].shape) == (H, W)


def test_batch_rejects_bad_pairs(dev, pairs):
    import torch
    from cvt_amd import _lib
    l, r, d = pairs[0]
    with pytest.raises(_lib.PmhError, match="depthmax"):
        dev.batch_depthmap([(l, r, d), (l, r, 0.0)], 5, 1)
    with pytest.raises(ValueError):
        dev.batch_depthmap([(l, r[:10].contiguous(), d)], 5, 1)
    with pytest.raises(_lib.PmhError, match="pd_iters"):
        dev.batch_depthmap([(l, r, d)], 5, 1, _lib.default_options(pd_iters=0))
    assert torch.isfinite(l).all()


def test_size_limits_are_errors(dev):
    """Rows wider than the row-scan kernels' LDS row (8192 px) are refused
    with an error before any launch, through both entry points; a window
    beyond half-size 64 likewise."""
    import torch
    from cvt_amd import _lib
    wide = torch.rand((2, 8193), device="cuda")
    with pytest.raises(_lib.PmhError, match="exceeds the supported 8192"):
        dev.depthmap(wide, wide, 1, 8.0, 1)
    with pytest.raises(_lib.PmhError, match="exceeds the supported 8192"):
        dev.batch_depthmap([(wide, wide, 8.0)], 1, 1)
    ok = torch.rand((2, 64), device="cuda")
    with pytest.raises(_lib.PmhError, match="half window"):
        dev.depthmap(ok, ok, 65, 8.0, 1)
    assert torch.isfinite(wide).all() and torch.isfinite(ok).all()
