"""Propagation parity at the benchmark window (35x35, half window 17) for every
launch shape the autotuner can pick.

The pass of pmhstereo_propagate_view (pmhstereo.cl:286-404) is run through
pmh_propagate with supplied draws, a live smoothness term (theta > 0) and a
view-propagation buffer, once with each propagate shape forced through a
tuning file: the one-wave kernel at LDS stages of 512, 704 and 896 texels and
the four-wave kernels (stage codes 1, 2, 3; 4 and 5 share one stage of
40-byte records per block).  The pair is 1100 px wide with
a disparity range of 900, so that a wave's span in the other view often
exceeds the LDS stage (15-22% of segment rows for the initial planes alone,
either view) and the overflow path (gathers through the L1) is taken.
Bars: every shape gives the same bits; against the oracle the planes are
identical except where the oracle's fold reaches the GPU's plane by ordering
costs tied within 1e-6 the other way; matching costs within 1e-5 relative.
"""
import numpy as np
import pytest

from proputil import compare_choice

pytestmark = pytest.mark.gpu

P = 17
SHAPES = ("2", "1", "3", "4", "5", "512", "704", "896")


def T(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def N(t):
    return t.detach().cpu().numpy()


def size_class(v):
    c = 0
    while (1 << c) < v:
        c += 1
    return c


def prop_key(W, H, sweep, lr):
    """The propagate tuning key of pmh_propagate (random_refine 2, six random
    neighbours)."""
    lanes = W if sweep == 0 else (W + 1) // 2
    seg = 62 if sweep == 0 else 63  # pixels per propagate tile (prop_seg_pixels)
    tiles = ((lanes + seg - 1) // seg) * H
    return f"propagate|p{P}|c1|s{sweep}|t{size_class(tiles)}|v{0 if lr else 1}|r2|n6"


def overflow_share(st, sweep, stage, W):
    """Share of (row segment, window row) pairs in which the span of the
    segment's own planes over the lanes' window rows exceeds `stage` texels
    (a lower bound on the overflow of the kernels' candidate unions)."""
    H = st.shape[0]
    hits = total = 0
    for y in range(H):
        xs = np.arange(W) if sweep == 0 else np.arange((y + sweep - 1) & 1, W, 2)
        for s0 in range(0, len(xs), 64):
            x = xs[s0:s0 + 64]
            a, b, c = (st[y, x, k].astype(np.float64) for k in range(3))
            for qy in range(max(0, y - P), min(H, y + P + 1)):
                lo = np.minimum(a * (x - P), a * (x + P)) + b * qy + c
                hi = np.maximum(a * (x - P), a * (x + P)) + b * qy + c
                lo, hi = np.clip(lo, 0, W - 1), np.clip(hi, 0, W - 1)
                hits += (hi.max() - lo.min() + 1) > stage
                total += 1
    return hits / total


@pytest.fixture(scope="module")
def wide(orc):
    from cvt_amd.synth import make_pair
    W, H, dmax = 1100, 12, 900.0
    pr = make_pair(W, H, 6, dmax, 0.01, seed=17)
    L, R = pr.left, pr.right
    return dict(W=W, H=H, dmax=dmax, L=L, R=R, GL=orc.gradxy(L), GR=orc.gradxy(R))


@pytest.fixture(scope="module")
def fresh_ctx():
    """A context of its own (contexts are per stream), autotuning off, so only
    the loaded tuning file chooses the shapes."""
    import torch
    from cvt_amd import ops
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        ctx = ops.context()
    ctx.autotune(False)
    return s, ctx


def _setup(orc, sc, lr, seed):
    rng = np.random.default_rng(seed)
    W, H, dmax = sc["W"], sc["H"], sc["dmax"]
    a = (sc["L"], sc["GL"], sc["R"], sc["GR"]) if lr else (sc["R"], sc["GR"], sc["L"], sc["GL"])
    old = orc.init(*a, P, dmax, lr)
    sm = np.zeros((H, W, 4), np.float32)
    sm[..., 0:2] = rng.uniform(-0.2, 0.2, (H, W, 2))
    sm[..., 2] = rng.uniform(0, 1, (H, W))
    sm[..., 3] = 1
    other = orc.init(a[2], a[3], a[0], a[1], P, dmax, 1 - lr)
    vb = orc.viewbuf_build(other)
    draws = rng.random((H, W, 21)).astype(np.float32)
    return a, old, sm, other, vb, draws


@pytest.mark.parametrize("lr,schedule", [(1, "redblack"), (0, "redblack"), (1, "jacobi"),
                                         (0, "jacobi")])
def test_propagate_p17_every_shape_matches_oracle(orc, wide, fresh_ctx, tmp_path, lr, schedule):
    import torch
    from cvt_amd import ops
    stream, ctx = fresh_ctx
    W, H, dmax, theta = wide["W"], wide["H"], wide["dmax"], 0.3
    a, old, sm, other, vb, draws = _setup(orc, wide, lr, 70 + lr)
    sweeps = (1, 2) if schedule == "redblack" else (0,)
    for s in sweeps:
        share = overflow_share(old, s, 896, W)
        print(f"lr={lr} sweep={s}: {share:.3f} of segment rows span more than 896 texels",
              flush=True)
        assert share > 0.05
    prev = old.copy()  # the black pass reads the red pixels written before it
    for s in sweeps:
        o = orc.propagate(old, *a, sm, theta, P, dmax, lr, 5, viewin=vb, viewsrc=other,
                          draws=draws, sweep=s, out=prev.copy())
        ref = None
        for shape in SHAPES:
            path = tmp_path / f"tune_{shape}.txt"
            path.write_text(f"{prop_key(W, H, s, lr)} {shape} 1\n")
            print(f"propagate P=17 {W}x{H} lr={lr} sweep={s} shape={shape}", flush=True)
            with torch.cuda.stream(stream):
                ctx.tuning_load(str(path))
                gout = T(prev)
                ops.propagate(T(old), *[T(x) for x in a], T(sm), theta, P, dmax, lr, 5,
                              viewin=T(vb), viewsrc=T(other), draws=T(draws), sweep=s, out=gout)
                g = N(gout)
            if ref is None:
                ref = g
                cctx = dict(old=old, out=prev, sm=sm, theta=theta, dmax=dmax, draws=draws, vb=vb,
                            vsrc=other, sweep=s)
                compare_choice(orc, a, g, o, P, lr, f"P=17 lr={lr} sweep={s} shape={shape}",
                               cctx)
            else:
                np.testing.assert_array_equal(g, ref, err_msg=f"shape {shape} vs {SHAPES[0]}")
        prev = ref
    with torch.cuda.stream(stream):
        saved = tmp_path / "saved.txt"
        ctx.tuning_save(str(saved))
    keys = [ln.split()[0] for ln in saved.read_text().splitlines() if ln.strip()]
    for s in sweeps:
        assert prop_key(W, H, s, lr) in keys  # the key format is the library's


@pytest.fixture(scope="module")
def wider(orc):
    from cvt_amd.synth import make_pair
    W, H, dmax = 2400, 8, 2000.0
    pr = make_pair(W, H, 6, dmax, 0.01, seed=18)
    L, R = pr.left, pr.right
    return dict(W=W, H=H, dmax=dmax, L=L, R=R, GL=orc.gradxy(L), GR=orc.gradxy(R))


@pytest.mark.parametrize("lr,schedule", [(1, "redblack"), (0, "redblack"), (0, "jacobi")])
def test_block_stage_overflow_matches_oracle(orc, wider, fresh_ctx, tmp_path, lr, schedule):
    """The block-stage kernels on a pair wider than their stage (1792 / 1152
    records): about a quarter of the segment rows' own planes span more than
    1792 texels, so their block unions overflow and take the per-wave quarter
    or gather paths, the shared neighbour planes' by gathers.  Same bars as
    above, against the one-wave kernel's bits and the oracle, with one tile per
    block and with persistent blocks."""
    import torch
    from cvt_amd import ops
    stream, ctx = fresh_ctx
    W, H, dmax, theta = wider["W"], wider["H"], wider["dmax"], 0.3
    a, old, sm, other, vb, draws = _setup(orc, wider, lr, 80 + lr)
    sweeps = (1, 2) if schedule == "redblack" else (0,)
    for s in sweeps:
        share = overflow_share(old, s, 1792, W)
        print(f"lr={lr} sweep={s}: {share:.3f} of segment rows span more than 1792 texels",
              flush=True)
        assert share > 0.1
    prev = old.copy()
    for s in sweeps:
        o = orc.propagate(old, *a, sm, theta, P, dmax, lr, 5, viewin=vb, viewsrc=other,
                          draws=draws, sweep=s, out=prev.copy())
        ref = None
        for shape, chunk in (("896", 1), ("4", 1), ("4", 0), ("5", 1), ("5", 0)):
            path = tmp_path / f"tune_{shape}_{chunk}.txt"
            path.write_text(f"{prop_key(W, H, s, lr)} {shape} {chunk}\n")
            print(f"propagate P=17 {W}x{H} lr={lr} sweep={s} shape={shape} chunk={chunk}",
                  flush=True)
            with torch.cuda.stream(stream):
                ctx.tuning_load(str(path))
                gout = T(prev)
                ops.propagate(T(old), *[T(x) for x in a], T(sm), theta, P, dmax, lr, 5,
                              viewin=T(vb), viewsrc=T(other), draws=T(draws), sweep=s, out=gout)
                g = N(gout)
            if ref is None:
                ref = g
                cctx = dict(old=old, out=prev, sm=sm, theta=theta, dmax=dmax, draws=draws, vb=vb,
                            vsrc=other, sweep=s)
                compare_choice(orc, a, g, o, P, lr, f"P=17 {W} px lr={lr} sweep={s}", cctx)
            else:
                np.testing.assert_array_equal(g, ref, err_msg=f"shape {shape}/{chunk} vs 896")
        prev = ref
