"""Pins of the multi-view restatement (oracle/pmh_oracle_mv.c; the reference's
draft is pmmvstereo.cl:97-304 with Plane.cl:60-86, which does not compile, so
there are no reference vectors): closed forms of the plane-induced homography,
projection of plane points computed independently in FP64, the bilinear cells
against the textbook interpolant, zero costs where a source view is an exact
copy or integer shift of the reference, an FP64 brute-force window cost, the
particle moves against their closed forms, a brute-force candidate fold, and
depth accuracy on a rendered scene."""
import numpy as np
import pytest

f32 = np.float32


def kmat(f, cx, cy):
    return np.array([[f, 0, cx], [0, f, cy], [0, 0, 1]], np.float64)


def rot(rng, amp):
    """Rotation by a random axis and angle <= amp (Rodrigues)."""
    a = rng.normal(size=3)
    a /= np.linalg.norm(a)
    th = rng.uniform(-amp, amp)
    A = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(th) * A + (1 - np.cos(th)) * A @ A


def tmat(R, t):
    return np.hstack([R, np.asarray(t, np.float64)[:, None]])


def particle(nx, ny, d):
    return np.array([nx, ny, d, 0], np.float32)


# ------------------------------------------------------------ homography ----
def test_identity_cameras_give_identity(orc):
    rng = np.random.default_rng(1)
    K = kmat(80.0, 40.0, 30.0)
    T = tmat(rot(rng, 0.2), rng.normal(size=3))
    kinv, M, e = orc.mv_constants(K, T, K[None], T[None])
    assert np.all(np.abs(e) < 1e-5)     # the float32 rotation is orthogonal to ~1e-7
    for _ in range(20):
        p = particle(rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5), rng.uniform(1, 20))
        np.testing.assert_allclose(orc.mv_homography(p, kinv, M[0], e[0]), np.eye(3),
                                   atol=1e-6)


@pytest.mark.parametrize("f,tx,ty,d", [(64.0, -0.25, 0.0, 4.0), (100.0, 0.3, -0.1, 7.5),
                                       (512.0, -0.05, 0.02, 3.0)])
def test_fronto_parallel_plane_is_a_shift(orc, f, tx, ty, d):
    """n = (0, 0, 1), source translated by t (no rotation): x' = x + f t / d."""
    K = kmat(f, 33.0, 21.0)
    Tr = tmat(np.eye(3), [0, 0, 0])
    Ts = tmat(np.eye(3), [tx, ty, 0])
    kinv, M, e = orc.mv_constants(K, Tr, K[None], Ts[None])
    H = orc.mv_homography(particle(0, 0, d), kinv, M[0], e[0])
    want = np.array([[1, 0, f * tx / d], [0, 1, f * ty / d], [0, 0, 1]])
    np.testing.assert_allclose(H, want, rtol=1e-6, atol=1e-5)


def test_homography_maps_plane_points(orc):
    """FP64, without the closed form: back-project a pixel onto the plane in
    the reference camera, move it to world and into the source camera."""
    rng = np.random.default_rng(2)
    for trial in range(10):
        Kr = kmat(rng.uniform(60, 900), rng.uniform(20, 400), rng.uniform(20, 300))
        Tr = tmat(rot(rng, 0.3), rng.normal(size=3))
        n = rng.integers(1, 5)
        Ks = np.stack([kmat(rng.uniform(60, 900), rng.uniform(20, 400), rng.uniform(20, 300))
                       for _ in range(n)])
        Ts = np.stack([tmat(rot(rng, 0.3), rng.normal(size=3)) for _ in range(n)])
        kinv, M, e = orc.mv_constants(Kr, Tr, Ks, Ts)
        p = particle(rng.uniform(-0.6, 0.6), rng.uniform(-0.6, 0.6), rng.uniform(2, 20))
        nv = np.array([p[0], p[1], np.sqrt(1 - float(p[0]) ** 2 - float(p[1]) ** 2)], np.float64)
        for v in range(n):
            H = orc.mv_homography(p, kinv, M[v], e[v]).astype(np.float64)
            for _ in range(5):
                x = np.array([rng.uniform(0, 400), rng.uniform(0, 300), 1.0])
                ray = np.linalg.solve(Kr, x)
                X = ray * (float(p[2]) / (nv @ ray))         # n . X = d in the reference camera
                Xw = Tr[:, :3].T @ (X - Tr[:, 3])
                q = Ks[v] @ (Ts[v][:, :3] @ Xw + Ts[v][:, 3])
                h = H @ x
                np.testing.assert_allclose(h[:2] / h[2], q[:2] / q[2], rtol=2e-4, atol=2e-3,
                                           err_msg=f"trial {trial} view {v}")


def test_depth_of_particle(orc):
    rng = np.random.default_rng(3)
    K = kmat(120.0, 60.0, 40.0)
    kinv, _, _ = orc.mv_constants(K, tmat(np.eye(3), [0, 0, 0]), K[None],
                                  tmat(np.eye(3), [0.1, 0, 0])[None])
    for _ in range(50):
        x, y = rng.uniform(0, 120), rng.uniform(0, 80)
        u = rng.uniform(0, 1, 3).astype(np.float32)
        p = orc.mv_particle_init(u, x, y, 1.0, 2.0, 12.0, kinv)
        z = 2.0 + float(u[0]) * 10.0
        assert abs(orc.mv_depth(p, kinv, x, y) - z) <= 2e-6 * z
    # a plane seen from its back side at this pixel: no positive depth
    p = particle(0.0, 0.0, -3.0)
    assert orc.mv_depth(p, kinv, 10.0, 10.0) == 0.0


# -------------------------------------------------------------- bilinear ----
def test_bilinear_cells_against_textbook(orc):
    rng = np.random.default_rng(4)
    img = rng.uniform(0, 1, (17, 23)).astype(np.float32)
    H, W = img.shape
    for j in range(H):
        for i in range(W):
            assert orc.mv_bilinear(img, float(i), float(j)) == img[j, i]
    for _ in range(300):
        x, y = rng.uniform(0, W), rng.uniform(0, H)
        x, y = float(f32(x)), float(f32(y))
        i0, j0 = int(x), int(y)
        i1, j1 = min(i0 + 1, W - 1), min(j0 + 1, H - 1)
        a, b = x - i0, y - j0
        I = img.astype(np.float64)
        want = ((1 - a) * (1 - b) * I[j0, i0] + a * (1 - b) * I[j0, i1] + (1 - a) * b * I[j1, i0]
                + a * b * I[j1, i1])
        assert abs(orc.mv_bilinear(img, x, y) - want) < 1e-6
    # a plane a + b x + c y is reproduced (up to rounding) everywhere inside
    yy, xx = np.mgrid[0:H, 0:W]
    ramp = (0.25 + 0.125 * xx + 0.0625 * yy).astype(np.float32)
    for _ in range(50):
        x, y = rng.uniform(0, W - 1), rng.uniform(0, H - 1)
        assert abs(orc.mv_bilinear(ramp, x, y) - (0.25 + 0.125 * x + 0.0625 * y)) < 1e-5
    cells = orc.mv_cells(img)
    assert np.all(cells[:, -1, 1] == 0) and np.all(cells[-1, :, 2] == 0)
    assert np.all(cells[:, -1, 3] == 0) and np.all(cells[-1, :, 3] == 0)


# ------------------------------------------------------------------ cost ----
def _views(orc, ref, srcs, K, Tr, Ts):
    return orc.MVViews(ref, K, Tr, list(srcs), np.repeat(K[None], len(srcs), 0), Ts)


def test_cost_zero_for_identical_views(orc):
    rng = np.random.default_rng(5)
    from cvt_amd.synth import _texture
    ref = _texture(rng, 24, 32, 1.25).astype(np.float32)
    K = kmat(40.0, 16.0, 12.0)
    Tr = tmat(np.eye(3), [0, 0, 0])
    v = _views(orc, ref, [ref, ref], K, Tr, np.stack([Tr, Tr]))
    for _ in range(20):
        p = particle(rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5), rng.uniform(1, 9))
        for (x, y) in [(16, 12), (3, 4), (30, 20)]:
            assert orc.mv_cost(v, x, y, p, 3) == 0.0


def test_cost_zero_for_integer_shift(orc):
    """Source = reference moved 4 px left; the fronto-parallel plane at the
    depth of a 4 px disparity maps every tap onto its own texel value."""
    rng = np.random.default_rng(6)
    from cvt_amd.synth import _texture
    W, H, k, P = 48, 20, 4, 3
    big = _texture(rng, H, W + k, 1.25).astype(np.float32)
    ref = big[:, k:].copy()                  # ref(x) = big(x + k)
    src = big[:, :W].copy()                  # src(x') = big(x'), so ref(x) = src(x + k)
    K = kmat(64.0, 24.0, 10.0)
    Tr = tmat(np.eye(3), [0, 0, 0])
    Ts = tmat(np.eye(3), [0.25, 0, 0])[None]  # x' = x + 64 * 0.25 / d = x + 4 at d = 4
    v = _views(orc, ref, [src], K, Tr, Ts)
    p = particle(0, 0, 4.0)
    H_ = orc.mv_homography(p, v.kinv, v.M[0], v.e[0])
    assert H_[0, 2] == 4.0
    for y in range(P + 1, H - P - 1):
        for x in range(P + 1, W - P - k - 2):
            assert orc.mv_cost(v, x, y, p, P) == 0.0, (x, y)
    # a plane 1 px off the true disparity costs more
    assert orc.mv_cost(v, 20, 10, particle(0, 0, 64 * 0.25 / 5), P) > 1e-3


def _cost_f64(v, x, y, p, P):
    """Brute-force FP64 restatement of the window cost (textbook bilinear,
    exact division, the homography from FP64 cameras)."""
    W, H = v.W, v.H
    n = np.array([p[0], p[1], np.sqrt(1 - float(p[0]) ** 2 - float(p[1]) ** 2)], np.float64)
    kinv = v.kinv.astype(np.float64).reshape(3, 3)
    m = kinv.T @ n / float(p[2])
    Ir, Gr = v.Ir.astype(np.float64), v.Gr.astype(np.float64)
    c0 = Ir[y, x]

    def bil(img, xp, yp):
        i0, j0 = int(xp), int(yp)
        i1, j1 = min(i0 + 1, W - 1), min(j0 + 1, H - 1)
        a, b = xp - i0, yp - j0
        return ((1 - a) * (1 - b) * img[j0, i0] + a * (1 - b) * img[j0, i1]
                + (1 - a) * b * img[j1, i0] + a * b * img[j1, i1])

    wsum = ret = 0.0
    for dy in range(-P, P + 1):
        for dx in range(-P, P + 1):
            qx, qy = x + dx, y + dy
            if not (0 <= qx < W and 0 <= qy < H):
                continue
            ray = kinv @ np.array([qx, qy, 1.0])
            if not float(p[2]) / (n @ ray) > 0:        # plane point behind the reference camera
                continue
            t = min(max(np.hypot(dx, dy) / 28.0, 0.0), 1.0)
            F = t * t * (3 - 2 * t) * 1.5 * 26 + 5
            w = np.exp(-3 * abs(c0 - Ir[qy, qx]) * F)
            for s in range(v.nsrc):
                Hm = v.M[s].astype(np.float64).reshape(3, 3) + np.outer(v.e[s], m)
                h = Hm @ np.array([qx, qy, 1.0])
                if not h[2] > 0:
                    continue
                xp, yp = h[0] / h[2], h[1] / h[2]
                if not (0 <= xp < W and 0 <= yp < H):
                    continue
                wsum += w
                Is = v.Is[s].astype(np.float64)
                Gs = v.Gs[s].astype(np.float64)
                cterm = 3 * min(abs(Ir[qy, qx] - bil(Is, xp, yp)), 0.04)
                gs = sum(min(abs(Gr[qy, qx, c] - bil(Gs[..., c], xp, yp)), 0.01)
                         for c in range(4))
                ret += w * (0.95 * gs + 0.05 * cterm)
    return 1e5 if wsum <= 1.1 else ret / wsum


@pytest.fixture(scope="module")
def mvscene(orc):
    from cvt_amd.synth import make_multiview
    s = make_multiview(64, 40, 2, seed=3)
    return s, orc.MVViews(s.ref, s.K, s.T, list(s.srcs), s.Ks, s.Ts)


def test_cost_against_f64_bruteforce(orc, mvscene):
    s, v = mvscene
    rng = np.random.default_rng(7)
    kinv = v.kinv
    for _ in range(25):
        x, y = int(rng.integers(0, v.W)), int(rng.integers(0, v.H))
        u = rng.uniform(0, 1, 3).astype(np.float32)
        # near the true depth half the time, anywhere otherwise
        if rng.uniform() < 0.5:
            z = s.depth[y, x] * rng.uniform(0.98, 1.02)
            u[0] = (z - s.dmin) / (s.dmax - s.dmin)
        p = orc.mv_particle_init(u, float(x), float(y), 1.0, s.dmin, s.dmax, kinv)
        got = orc.mv_cost(v, x, y, p, 4)
        want = _cost_f64(v, x, y, p, 4)
        assert abs(got - want) <= 1e-4 * max(abs(want), 1e-3), (x, y, got, want)


def test_cost_rejects_bad_particles(orc, mvscene):
    s, v = mvscene
    assert orc.mv_cost(v, 10, 10, particle(np.nan, 0, 5), 3) == 1e5
    assert orc.mv_cost(v, 10, 10, particle(0, 0, np.inf), 3) == 1e5
    # a plane behind both source cameras: no tap lands, support weight 0
    assert orc.mv_cost(v, 10, 10, particle(0, 0, -5.0), 3) == 1e5
    # half window 0: a single tap has weight at most 2 (two views), 1 <= 1.1 for one view
    v1 = orc.MVViews(s.ref, s.K, s.T, [s.srcs[0]], s.Ks[:1], s.Ts[:1])
    z = float(s.depth[20, 30])
    p = orc.mv_particle_init(np.array([(z - 3) / 11, 0.5, 0.5], np.float32), 30.0, 20.0, 1.0,
                             3.0, 14.0, v.kinv)
    assert orc.mv_cost(v1, 30, 20, p, 0) == 1e5


# ------------------------------------------------------------ particles ----
@pytest.mark.parametrize("u", [(0.3, 0.5, 0.5), (0.9, 0.1, 0.8), (0.0, 1.0, 0.0), (0.5, 0.2, 0.93)])
def test_particle_init_matches_stereo_normal(orc, u):
    """The normal is nd_state_init's (pmhstereo.cl:50-68), read back through the
    stereo oracle's independent plane algebra; the depth at the pixel is
    dmin + u0 (dmax - dmin)."""
    K = kmat(100.0, 50.0, 40.0)
    kinv, _, _ = orc.mv_constants(K, tmat(np.eye(3), [0, 0, 0]), K[None],
                                  tmat(np.eye(3), [0.1, 0, 0])[None])
    x, y = 37.0, 11.0
    for nm in (1.0, 2.0):
        p = orc.mv_particle_init(np.array(u, np.float32), x, y, nm, 2.0, 12.0, kinv)
        nd = orc.state_to_ref_normal_depth(orc.state_init(np.array(u, np.float32), x, y, 1, nm,
                                                          64.0), x, y, 1)
        np.testing.assert_allclose(p[:2], nd[:2], atol=2e-6)
        z = 2.0 + u[0] * 10.0
        assert abs(orc.mv_depth(p, kinv, x, y) - z) <= 2e-6 * z


def test_particle_refine_closed_form(orc):
    rng = np.random.default_rng(8)
    K = kmat(100.0, 50.0, 40.0)
    kinv, _, _ = orc.mv_constants(K, tmat(np.eye(3), [0, 0, 0]), K[None],
                                  tmat(np.eye(3), [0.1, 0, 0])[None])
    dmin, dmax = 2.0, 12.0
    for _ in range(100):
        x, y = rng.uniform(0, 100), rng.uniform(0, 80)
        p = orc.mv_particle_init(rng.uniform(0, 1, 3).astype(np.float32), x, y, 1.0, dmin, dmax,
                                 kinv)
        u = rng.uniform(0, 1, 3).astype(np.float32)
        r = orc.mv_particle_refine(u, p, x, y, dmin, dmax, kinv)
        z0 = orc.mv_depth(p, kinv, x, y)
        z = np.clip(z0 + (u[0] - 0.5) * (dmax - dmin) / 32, dmin, dmax)
        assert abs(orc.mv_depth(r, kinv, x, y) - z) <= 1e-5 * z
        nx = np.clip(float(p[0]) + (float(u[1]) - 0.5) * 0.2, -0.95, 0.95)
        ny = np.clip(float(p[1]) + (float(u[2]) - 0.5) * 0.2, -0.95, 0.95)
        nf = max(np.hypot(nx, ny) + 0.001, 1.0)
        np.testing.assert_allclose(r[:2], [nx / nf, ny / nf], atol=1e-6)


# ------------------------------------------------------------- propagate ----
def _fold(orc, v, old, out, x, y, P, dmin, dmax, u, sweep, nref):
    """pmmv3_iteration (pmmvstereo.cl:237-296) for one pixel, enumerated."""
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
    self_ = old[y, x].copy()
    for c in cands:
        c = c.copy()
        c[3] = orc.mv_cost(v, x, y, c, P)
        if c[3] < self_[3]:
            self_ = c
    for i in range(nref):
        c = orc.mv_particle_refine(u[9 + 3 * i:12 + 3 * i], self_, float(x), float(y), dmin,
                                   dmax, v.kinv)
        c[3] = orc.mv_cost(v, x, y, c, P)
        if c[3] < self_[3]:
            self_ = c
    return self_


@pytest.mark.parametrize("nref", [0, 2])
def test_propagate_against_enumeration(orc, mvscene, nref):
    s, v = mvscene
    P, it = 2, 3
    sub = orc.MVViews(s.ref[:14, :20], s.K, s.T, [im[:14, :20] for im in s.srcs], s.Ks, s.Ts)
    old = orc.mv_init(sub, P, s.dmin, s.dmax)
    H, W = old.shape[:2]
    draws = np.stack([np.stack([orc.rng_draws(y * W + x + it, 33, 33) for x in range(W)])
                      for y in range(H)])
    for sweep in (0, 1, 2):
        out = old.copy() if sweep != 2 else orc.mv_propagate(sub, old, P, s.dmin, s.dmax, it,
                                                              draws, 1, refine=nref)
        res = orc.mv_propagate(sub, old, P, s.dmin, s.dmax, it, draws, sweep, out.copy(),
                               refine=nref)
        for y in range(H):
            for x in range(W):
                if sweep and (x + y) % 2 != sweep - 1:
                    assert np.array_equal(res[y, x], out[y, x])
                    continue
                want = _fold(orc, sub, old, out, x, y, P, s.dmin, s.dmax, draws[y, x], sweep,
                             nref)
                assert np.array_equal(res[y, x], want), (sweep, x, y)
        # the MWC64X streams of pmmvstereo.cl:237 are the supplied draws
        alt = orc.mv_propagate(sub, old, P, s.dmin, s.dmax, it, None, sweep, out.copy(),
                               refine=nref)
        assert np.array_equal(alt, res)


def test_init_uses_its_draws(orc, mvscene):
    s, v = mvscene
    out = orc.mv_init(v, 3, s.dmin, s.dmax)
    for (x, y) in [(0, 0), (5, 7), (63, 39)]:
        u = orc.rng_draws(y * v.W + x, 3, 3)
        p = orc.mv_particle_init(u, float(x), float(y), 1.0, s.dmin, s.dmax, v.kinv)
        p[3] = orc.mv_cost(v, x, y, p, 3)
        assert np.array_equal(out[y, x], p)


# ------------------------------------------------------------- depth map ----
def test_depthmap_is_the_composition(orc, mvscene):
    s, v = mvscene
    P, iters = 3, 2
    for rb in (False, True):
        d, parts = orc.mv_depthmap(v, P, s.dmin, s.dmax, iters, red_black=rb, refine=2)
        a = orc.mv_init(v, P, s.dmin, s.dmax)
        for it in range(iters):
            if rb:
                b = orc.mv_propagate(v, a, P, s.dmin, s.dmax, it, None, 1, a.copy())
                b = orc.mv_propagate(v, a, P, s.dmin, s.dmax, it, None, 2, b)
            else:
                b = orc.mv_propagate(v, a, P, s.dmin, s.dmax, it, None, 0, a.copy())
            a = b
        assert np.array_equal(parts, a)
        want = np.array([[orc.mv_depth(a[y, x], v.kinv, float(x), float(y))
                          for x in range(v.W)] for y in range(v.H)], np.float32)
        assert np.array_equal(d, want)


def test_depthmap_recovers_scene_depth(orc):
    from cvt_amd.synth import make_multiview, depth_error_rate
    s = make_multiview(160, 96, 2, seed=3)
    v = orc.MVViews(s.ref, s.K, s.T, list(s.srcs), s.Ks, s.Ts)
    d, _ = orc.mv_depthmap(v, 4, s.dmin, s.dmax, 4, red_black=True, refine=2)
    assert depth_error_rate(d, s, 0.05) < 6.0
    # a wrong sign of the plane distance would put no depth near the truth
    d0 = orc.mv_init(v, 4, s.dmin, s.dmax)
    init_depth = np.array([[orc.mv_depth(d0[y, x], v.kinv, float(x), float(y))
                            for x in range(v.W)] for y in range(v.H)])
    assert depth_error_rate(init_depth, s, 0.05) > 60.0
