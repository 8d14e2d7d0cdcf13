"""Brute-force restatement of one pixel of pmmv3_iteration (pmmvstereo.cl:
236-296) for the multi-view parity tests: the candidates in the draft's order,
costed with the oracle's window cost and folded with strict `<`.  `outcomes`
follows both branches of every comparison whose two costs are within a tie
tolerance, giving every particle the algorithm can end on when near-equal
costs are ordered either way."""
import numpy as np

f32 = np.float32


def candidates(orc, v, x, y, old, out, dmin, dmax, u, sweep):
    H, W = old.shape[:2]

    def cur(nx, ny):
        nx, ny = min(max(nx, 0), W - 1), min(max(ny, 0), H - 1)
        src = out if sweep == 2 and (nx + ny) % 2 == 0 else old
        return src[ny, nx].copy()

    cands = [cur(x + px, y + py) for py in (-1, 0, 1) for px in (-1, 0, 1) if (px, py) != (0, 0)]
    for i in range(3):
        cands.append(cur(x + int(f32(u[2 * i] * f32(7)) + f32(0.5)),
                         y + int(f32(u[2 * i + 1] * f32(7)) + f32(0.5))))
    cands.append(orc.mv_particle_init(u[6:9], float(x), float(y), 2.0, dmin, dmax, v.kinv))
    return cands


def outcomes(orc, v, x, y, old, out, P, dmin, dmax, u, sweep, nref, tol=2e-6, limit=64):
    costed = []
    for c in candidates(orc, v, x, y, old, out, dmin, dmax, u, sweep):
        c = c.copy()
        c[3] = orc.mv_cost(v, x, y, c, P)
        costed.append(c)
    results = set()

    def tie(a, b):
        return abs(float(a) - float(b)) <= tol * max(1e-3, abs(float(b)))

    def fold(k, s):
        if len(results) > limit:
            return
        if k < len(costed):
            c = costed[k]
            if tie(c[3], s[3]):
                fold(k + 1, c)
                fold(k + 1, s)
            else:
                fold(k + 1, c if c[3] < s[3] else s)
            return
        refine(0, s)

    def refine(i, s):
        if i == nref:
            results.add(tuple(np.asarray(s[:3], np.float32).tolist()))
            return
        c = orc.mv_particle_refine(u[9 + 3 * i:12 + 3 * i], s, float(x), float(y), dmin, dmax,
                                   v.kinv)
        c[3] = orc.mv_cost(v, x, y, c, P)
        if tie(c[3], s[3]):
            refine(i + 1, c)
            refine(i + 1, s)
        else:
            refine(i + 1, c if c[3] < s[3] else s)

    fold(0, old[y, x].copy())
    return results


def compare_choice(orc, v, g, o, old, out, P, dmin, dmax, draws, sweep, nref, what,
                   rtol_cost=1e-5, max_diff=0.01):
    """Particles g (GPU) and o (oracle) of one pass must match, except where
    the oracle's fold reaches the GPU's particle by ordering near-tied costs
    the other way; costs of matching particles within rtol_cost."""
    diff = np.any(g[..., :3] != o[..., :3], axis=-1)
    if sweep != 0:
        par = np.add.outer(np.arange(g.shape[0]), np.arange(g.shape[1])) & 1
        diff &= par == (0 if sweep == 1 else 1)
    bad = []
    for y, x in zip(*np.nonzero(diff)):
        outs = outcomes(orc, v, int(x), int(y), old, out, P, dmin, dmax, draws[y, x], sweep,
                        nref)
        if tuple(g[y, x, :3].tolist()) not in outs:
            bad.append((int(x), int(y), len(outs)))
    assert not bad, f"{what}: {len(bad)} mismatches not explained by ties, e.g. {bad[:3]}"
    assert diff.mean() < max_diff, f"{what}: {diff.mean():.4f} of pixels differ"
    np.testing.assert_allclose(g[~diff][:, 3], o[~diff][:, 3], rtol=rtol_cost, atol=1e-7)
