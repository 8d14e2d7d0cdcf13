"""Brute-force restatement of one pixel of pmhstereo_propagate_view
(pmhstereo.cl:320-389) for the parity tests: candidates enumerated in the
reference order, costed with the oracle's window cost, folded with `<=` in
float32.  `outcomes` also follows both branches of every acceptance test
whose two sides are within a tie tolerance, giving every plane the reference
algorithm can end on when near-equal costs are ordered either way."""
import numpy as np

f32 = np.float32


def smooth_norm(sp, dmax):
    sx, sy = f32(sp[0]), f32(sp[1])
    nf = max(f32(np.sqrt(f32(sx * sx + sy * sy))) + f32(0.001), f32(1.0))
    sx, sy = f32(sx / nf), f32(sy / nf)
    sz = f32(np.sqrt(f32(f32(f32(1) - f32(sx * sx)) - f32(sy * sy))))
    return np.array([sx, sy, sz, f32(sp[2]) * f32(dmax)], np.float32)


def sdist(orc, s, sm, cx, cy, dmax, lr):
    A = orc.state_to_ref_normal_depth(s, cx, cy, lr)
    d = (A - sm).astype(np.float32)
    dw = f32(1.0) / f32(f32(dmax) * f32(dmax))
    return f32(f32(f32(d[0] * d[0] + d[1] * d[1]) + d[2] * f32(f32(0) * d[2])) + d[3] * f32(dw * d[3]))


def sides(orc, c, s, sm, cx, cy, dmax, lr, theta):
    fin = np.all(np.isfinite(c[:3])) and np.all(np.isfinite(s[:3]))
    a = sdist(orc, c, sm, cx, cy, dmax, lr) if fin else f32(0)
    b = sdist(orc, s, sm, cx, cy, dmax, lr) if fin else f32(0)
    return f32(c[3] + f32(f32(theta) * a)), f32(s[3] + f32(f32(theta) * b))


def accept(orc, c, s, sm, cx, cy, dmax, lr, theta):
    lhs, rhs = sides(orc, c, s, sm, cx, cy, dmax, lr, theta)
    return lhs <= rhs


def candidates(orc, x, y, old, out, sm_img, dmax, lr, draws, vb, vsrc, sweep):
    """The 16 + n_view candidates before refinement, in the reference order."""
    H, W = old.shape[:2]

    def cur(nx, ny):
        nx, ny = min(max(nx, 0), W - 1), min(max(ny, 0), H - 1)
        if sweep == 2 and ((nx + ny) & 1) == 0:
            return out[ny, nx].copy()
        return old[ny, nx].copy()

    cx, cy = f32(x), f32(y)
    u = draws[y, x]
    sm = smooth_norm(sm_img[y, x], dmax)
    cands = [cur(x + px, y + py) for py in (-1, 0, 1) for px in (-1, 0, 1) if (px, py) != (0, 0)]
    s4 = np.array([f32(1) - f32(-sm[0] / sm[2]), f32(0) - f32(-sm[1] / sm[2]),
                   f32(0) - f32(f32(f32(f32(sm[0] * cx) + f32(sm[1] * cy)) / sm[2]) + sm[3]), 0],
                  np.float32)
    cands.append(s4 if lr else orc.state_viewprop(s4))
    for i in range(6):
        ox = int(f32(u[2 * i] * f32(7)) + f32(0.5))
        oy = int(f32(u[2 * i + 1] * f32(7)) + f32(0.5))
        cands.append(cur(x + ox, y + oy))
    cands.append(orc.state_init(u[12:15], cx, cy, lr, 2.0, dmax))
    if vb is not None:
        n = min(int(vb[y, x, 0]), 4)
        for i in range(n):
            cands.append(orc.state_viewprop(vsrc[y, int(vb[y, x, 1 + i])]))
    return cands, sm, u


def outcomes(orc, imgs, x, y, old, out, sm_img, theta, P, dmax, lr, draws, vb, vsrc, sweep,
             tol=1e-6, limit=64):
    """Every final plane (a, b, c) of the fold when near-ties go either way."""
    cands, sm, u = candidates(orc, x, y, old, out, sm_img, dmax, lr, draws, vb, vsrc, sweep)
    cx, cy = f32(x), f32(y)
    costed = []
    for c in cands:
        c = c.copy()
        c[3] = orc.cost(*imgs, x, y, c, P)
        costed.append(c)
    results = set()

    def tie(lhs, rhs):
        return abs(float(lhs) - float(rhs)) <= tol * max(1.0, abs(float(rhs)))

    def fold(k, self_):
        if len(results) > limit:
            return
        if k < len(costed):
            c = costed[k]
            lhs, rhs = sides(orc, c, self_, sm, cx, cy, dmax, lr, theta)
            if tie(lhs, rhs):
                fold(k + 1, c)
                fold(k + 1, self_)
            else:
                fold(k + 1, c if lhs <= rhs else self_)
            return
        refine(0, self_)

    def refine(i, self_):
        if i == 2:
            results.add(tuple(np.asarray(self_[:3], np.float32).tolist()))
            return
        c = orc.state_refine(u[15 + 3 * i:18 + 3 * i], self_, cx, cy, dmax, lr)
        c[3] = orc.cost(*imgs, x, y, c, P)
        lhs, rhs = sides(orc, c, self_, sm, cx, cy, dmax, lr, theta)
        if tie(lhs, rhs):
            refine(i + 1, c)
            refine(i + 1, self_)
        else:
            refine(i + 1, c if lhs <= rhs else self_)

    fold(0, old[y, x].copy())
    return results


def compare_choice(orc, a, g, o, P, lr, what, ctx, tie=1e-6, rtol_cost=1e-5, max_diff=0.01):
    """Planes g (GPU) and o (oracle) must match, except where the oracle's fold
    reaches the GPU's plane by ordering near-tied costs (within `tie`) the
    other way; costs of matching planes within rtol_cost.  ctx: the pass's
    inputs (old, out, sm, theta, dmax, draws, vb, vsrc, sweep)."""
    diff = np.any(g[..., :3] != o[..., :3], axis=-1)
    if ctx["sweep"] != 0:
        par = (np.add.outer(np.arange(g.shape[0]), np.arange(g.shape[1])) & 1)
        diff &= par == (0 if ctx["sweep"] == 1 else 1)
    bad = []
    for y, x in zip(*np.nonzero(diff)):
        outs = outcomes(orc, a, int(x), int(y), ctx["old"], ctx["out"], ctx["sm"], ctx["theta"],
                        P, ctx["dmax"], lr, ctx["draws"], ctx["vb"], ctx["vsrc"], ctx["sweep"],
                        tol=tie)
        if tuple(g[y, x, :3].tolist()) not in outs:
            bad.append((int(x), int(y), len(outs)))
    assert not bad, f"{what}: {len(bad)} mismatches not explained by ties, e.g. {bad[:3]}"
    assert diff.mean() < max_diff, f"{what}: {diff.mean():.4f} of pixels differ"
    np.testing.assert_allclose(g[~diff][:, 3], o[~diff][:, 3], rtol=rtol_cost, atol=1e-7)
