"""Seeded synthetic rectified stereo pairs with ground truth, and bad-pixel scoring.

The reference ships no stereo data with ground truth, so the workloads are
generated: ``n_planes`` random slanted planes in Voronoi regions, a
band-limited noise texture in [0, 1] on the left image, the right image warped
from the left (forward splat with a z-buffer, nearer surface wins), holes
filled with an independent texture, plus Gaussian noise.  Disparity convention
of the reference plane field (pmhstereo.cl:50-68): left pixel x matches right
pixel x - d.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy.ndimage import gaussian_filter


@dataclass
class StereoPair:
    left: np.ndarray        # (H, W) float32, or (H, W, 4) RGBA (alpha 1)
    right: np.ndarray       # like left
    disp: np.ndarray        # (H, W) float32, left-view ground-truth disparity (px)
    nonocc: np.ndarray      # (H, W) bool, left pixels visible in the right view
    dmax: float
    seed: int

    @property
    def width(self) -> int:
        return self.left.shape[1]

    @property
    def height(self) -> int:
        return self.left.shape[0]


def _texture(rng: np.random.Generator, H: int, W: int, blur: float) -> np.ndarray:
    t = gaussian_filter(rng.standard_normal((H, W)).astype(np.float64), blur, mode="reflect")
    lo, hi = np.percentile(t, 0.5), np.percentile(t, 99.5)
    return np.clip((t - lo) / (hi - lo), 0.0, 1.0)


def make_pair(width: int, height: int, n_planes: int, dmax: float, noise: float = 0.01,
              seed: int = 1, blur: float = 1.25, colour: bool = False) -> StereoPair:
    """A seeded pair; ``colour`` gives RGBA images (three independent textures,
    alpha 1) over the same scene geometry."""
    rng = np.random.default_rng(seed)
    W, H = int(width), int(height)
    sx = rng.uniform(0, W, n_planes)
    sy = rng.uniform(0, H, n_planes)
    d0 = rng.uniform(0.1, 0.9, n_planes) * dmax
    smax = min(0.3, 0.4 * dmax / np.sqrt(W * H / n_planes))
    gx = rng.uniform(-smax, smax, n_planes)
    gy = rng.uniform(-smax, smax, n_planes)

    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    # Voronoi labels, computed in row blocks to bound memory
    lab = np.empty((H, W), np.int32)
    for r0 in range(0, H, 256):
        r1 = min(H, r0 + 256)
        d2 = (xx[r0:r1, :, None] - sx) ** 2 + (yy[r0:r1, :, None] - sy) ** 2
        lab[r0:r1] = np.argmin(d2, axis=2)
    disp = d0[lab] + gx[lab] * (xx - sx[lab]) + gy[lab] * (yy - sy[lab])
    disp = np.clip(disp, 0.0, dmax)

    C = 3 if colour else 1
    left = np.stack([_texture(rng, H, W, blur) for _ in range(C)], -1)
    hole_tex = np.stack([_texture(rng, H, W, blur) for _ in range(C)], -1)

    # forward splat along each row: segment (x, x+1) maps to [x - d(x), x+1 - d(x+1)]
    xr0 = xx[:, :-1] - disp[:, :-1]
    xr1 = xx[:, 1:] - disp[:, 1:]
    same = lab[:, :-1] == lab[:, 1:]
    span = xr1 - xr0
    ok = same & (span > 1e-6)
    cand_pix, cand_val, cand_dep = [], [], []
    for k in range(3):
        xi = np.ceil(xr0) + k
        sel = ok & (xi <= xr1) & (xi >= 0) & (xi < W)
        t = np.where(sel, (xi - xr0) / np.where(span > 0, span, 1.0), 0.0)
        r, c = np.nonzero(sel)
        tt = t[r, c]
        cand_pix.append(r.astype(np.int64) * W + xi[r, c].astype(np.int64))
        cand_val.append((1 - tt)[:, None] * left[r, c] + tt[:, None] * left[r, c + 1])
        cand_dep.append((1 - tt) * disp[r, c] + tt * disp[r, c + 1])
    pix = np.concatenate(cand_pix)
    val = np.concatenate(cand_val)
    dep = np.concatenate(cand_dep)
    order = np.lexsort((dep, pix))           # by pixel, then depth ascending
    pix, val, dep = pix[order], val[order], dep[order]
    last = np.ones(pix.shape, bool)
    last[:-1] = pix[1:] != pix[:-1]          # keep the largest disparity per pixel
    right = hole_tex.reshape(-1, C).copy()
    rdep = np.full(H * W, -1.0)
    right[pix[last]] = val[last]
    rdep[pix[last]] = dep[last]
    right = right.reshape(H, W, C)
    rdep = rdep.reshape(H, W)

    # left pixel visible iff the right z-buffer at its match holds its surface
    xm = np.rint(xx - disp).astype(np.int64)
    inside = (xm >= 0) & (xm < W)
    xm_c = np.clip(xm, 0, W - 1)
    zb = rdep[np.arange(H)[:, None], xm_c]
    nonocc = inside & (zb - disp < 1.0)

    right = right + rng.normal(0.0, noise, (H, W, C))
    if colour:
        alpha = np.ones((H, W, 1))
        left = np.concatenate([left, alpha], -1)
        right = np.concatenate([right, alpha], -1)
    else:
        left, right = left[..., 0], right[..., 0]
    return StereoPair(left.astype(np.float32), right.astype(np.float32),
                      disp.astype(np.float32), nonocc, float(dmax), int(seed))


def bad_pixels(disp_est: np.ndarray, pair: StereoPair, thr: float = 1.0) -> float:
    """Percentage of non-occluded ground-truth pixels with |d - d_gt| > thr."""
    m = pair.nonocc
    err = np.abs(np.asarray(disp_est, np.float64)[m] - pair.disp[m])
    err[~np.isfinite(err)] = np.inf
    return float(100.0 * np.mean(err > thr))


# Workloads named by what they are (sizes, planes, disparity range, noise).
def pair_450x375(seed: int = 1) -> StereoPair:
    return make_pair(450, 375, 8, 64.0, 0.01, seed)


def pair_1242x375(seed: int = 2) -> StereoPair:
    return make_pair(1242, 375, 16, 192.0, 0.01, seed)


def pair_2880x1988(seed: int = 5) -> StereoPair:
    return make_pair(2880, 1988, 32, 400.0, 0.02, seed)


def _make_many(args, workers=None):
    """make_pair over argument tuples in threads (numpy releases the GIL in the
    heavy steps); every pair has its own seed, so the result does not depend
    on the thread count."""
    from concurrent.futures import ThreadPoolExecutor
    import os
    workers = workers or min(8, os.cpu_count() or 1)
    with ThreadPoolExecutor(workers) as ex:
        return list(ex.map(lambda a: make_pair(*a), args))


def uneven_batch(n: int = 64, seed: int = 4, workers: int | None = None):
    """n pairs with W in [320,1920], H in [240,1080], dmax in [32,256], 4-24 planes."""
    rng = np.random.default_rng(seed)
    args = []
    for i in range(n):
        W = int(rng.integers(320, 1921))
        H = int(rng.integers(240, 1081))
        dm = float(rng.integers(32, 257))
        npl = int(rng.integers(4, 25))
        args.append((W, H, npl, dm, 0.01, seed * 1000 + i))
    return _make_many(args, workers)


def large_batch(n: int = 8, seed: int = 5, workers: int | None = None):
    """n pairs of 2880x1988, 32 planes, disparity 0-400, noise 0.02 (seeds
    seed, seed + 1, ...)."""
    return _make_many([(2880, 1988, 32, 400.0, 0.02, seed + k) for k in range(n)], workers)


# ------------------------------------------------------------ multi-view ----
@dataclass
class MultiViewScene:
    """A reference view and nsrc source views of one seeded piecewise-planar
    scene, with the reference view's ground-truth depth.  Cameras: K (3x3) and
    T = [R | t] (3x4, world -> camera); the world frame is the reference
    camera's (x right, y down, z forward)."""
    ref: np.ndarray         # (H, W) float32 gray
    srcs: np.ndarray        # (nsrc, H, W) float32
    K: np.ndarray           # (3, 3) float32, all views
    T: np.ndarray           # (3, 4) float32, reference: [I | 0]
    Ks: np.ndarray          # (nsrc, 3, 3)
    Ts: np.ndarray          # (nsrc, 3, 4)
    depth: np.ndarray       # (H, W) float32 ground truth (z in the reference camera)
    visible: np.ndarray     # (H, W) bool: the surface point is seen by a source view
    dmin: float
    dmax: float
    seed: int

    @property
    def width(self) -> int:
        return self.ref.shape[1]

    @property
    def height(self) -> int:
        return self.ref.shape[0]


def _look_at(C, target):
    z = target - C
    z = z / np.linalg.norm(z)
    x = np.cross(np.array([0.0, 1.0, 0.0]), z)
    x = x / np.linalg.norm(x)
    y = np.cross(z, x)
    R = np.stack([x, y, z])
    return np.hstack([R, (-R @ C)[:, None]])


def _render(surfaces, K, T, W, H):
    """Ray-casts the surfaces into one view: (value, z in that camera)."""
    from scipy.ndimage import map_coordinates
    R, t = T[:, :3], T[:, 3]
    C = -R.T @ t
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    rays = np.stack([xx, yy, np.ones_like(xx)], -1) @ np.linalg.inv(K).T   # camera frame
    dirs = rays @ R                                                         # world frame
    best = np.full((H, W), np.inf)
    val = np.zeros((H, W))
    for s in surfaces:
        den = dirs @ s["n"]
        with np.errstate(divide="ignore", invalid="ignore"):
            lam = (s["n"] @ (s["p"] - C)) / den
        X = C + lam[..., None] * dirs
        rel = X - s["p"]
        u = rel @ s["a1"]
        v = rel @ s["a2"]
        hit = (lam > 1e-6) & np.isfinite(lam) & (np.abs(u) <= s["h1"]) & (np.abs(v) <= s["h2"])
        nearer = hit & (lam < best)
        if not nearer.any():
            continue
        tu = u[nearer] * s["scale"] + s["tex"].shape[1] / 2
        tv = v[nearer] * s["scale"] + s["tex"].shape[0] / 2
        val[nearer] = map_coordinates(s["tex"], [tv, tu], order=1, mode="wrap")
        best[nearer] = lam[nearer]
    # z in this camera = lam * (R dir).z = lam (rays have z = 1)
    return val, best


def make_multiview(width: int, height: int, nsrc: int = 2, seed: int = 1, n_boxes: int = 4,
                   noise: float = 0.01, blur: float = 1.25) -> MultiViewScene:
    """A seeded scene: a slanted textured background plane at depth 10-12 and
    n_boxes textured slanted rectangles at depth 4-8, seen by a reference
    camera at the origin and nsrc source cameras on a ring of radius 0.5
    around it (looking at the scene centre), focal length 0.8 W.  Depth range
    of the scene: dmin 3, dmax 14."""
    if not 1 <= nsrc <= 4:
        raise ValueError("1 to 4 source views")
    rng = np.random.default_rng(seed)
    W, H = int(width), int(height)
    f = 0.8 * W
    K = np.array([[f, 0.0, (W - 1) / 2.0], [0.0, f, (H - 1) / 2.0], [0.0, 0.0, 1.0]])
    Tr = np.hstack([np.eye(3), np.zeros((3, 1))])
    scale = f / 6.0

    def surface(p, n, h1, h2):
        n = n / np.linalg.norm(n)
        a1 = np.cross(np.array([0.0, 1.0, 0.0]), n)
        a1 /= np.linalg.norm(a1)
        a2 = np.cross(n, a1)
        side = int(2 * max(h1, h2) * scale) + 64 if np.isfinite(max(h1, h2)) else 2048
        side = min(side, 2048)
        return dict(p=p, n=n, a1=a1, a2=a2, h1=h1, h2=h2, scale=scale,
                    tex=_texture(rng, side, side, blur))

    zb = rng.uniform(10.0, 12.0)
    surfaces = [surface(np.array([0.0, 0.0, zb]),
                        np.array([rng.uniform(-0.25, 0.25), rng.uniform(-0.25, 0.25), -1.0]),
                        np.inf, np.inf)]
    for _ in range(n_boxes):
        z0 = rng.uniform(4.0, 8.0)
        x0 = rng.uniform(-0.35, 0.35) * W / f * z0
        y0 = rng.uniform(-0.35, 0.35) * H / f * z0
        n = np.array([rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5), -1.0])
        surfaces.append(surface(np.array([x0, y0, z0]), n, rng.uniform(0.08, 0.2) * W / f * z0,
                                rng.uniform(0.1, 0.3) * H / f * z0))

    target = np.array([0.0, 0.0, 8.0])
    Ts = []
    for k in range(nsrc):
        ang = 2 * np.pi * k / nsrc + rng.uniform(-0.3, 0.3)
        C = np.array([0.5 * np.cos(ang), 0.3 * np.sin(ang), rng.uniform(-0.1, 0.1)])
        Ts.append(_look_at(C, target))
    ref, zref = _render(surfaces, K, Tr, W, H)
    srcs, zs = [], []
    for T in Ts:
        v, z = _render(surfaces, K, T, W, H)
        srcs.append(v)
        zs.append(z)

    # a reference pixel is visible if some source sees its surface point unoccluded
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    Xw = (np.stack([xx, yy, np.ones_like(xx)], -1) @ np.linalg.inv(K).T) * zref[..., None]
    visible = np.zeros((H, W), bool)
    for T, z in zip(Ts, zs):
        Xc = Xw @ T[:, :3].T + T[:, 3]
        with np.errstate(divide="ignore", invalid="ignore"):
            px = (Xc @ K.T)
            u, v = px[..., 0] / px[..., 2], px[..., 1] / px[..., 2]
        ok = np.isfinite(u) & np.isfinite(v) & (u >= 0) & (u <= W - 1) & (v >= 0) & (v <= H - 1)
        ui = np.clip(np.rint(np.where(ok, u, 0)).astype(np.int64), 0, W - 1)
        vi = np.clip(np.rint(np.where(ok, v, 0)).astype(np.int64), 0, H - 1)
        visible |= ok & (np.abs(z[vi, ui] - Xc[..., 2]) < 0.01 * Xc[..., 2])

    ref = ref + rng.normal(0.0, noise, ref.shape)
    srcs = np.stack(srcs) + rng.normal(0.0, noise, (nsrc, H, W))
    f32 = np.float32
    return MultiViewScene(ref.astype(f32), srcs.astype(f32), K.astype(f32), Tr.astype(f32),
                          np.repeat(K[None].astype(f32), nsrc, 0), np.stack(Ts).astype(f32),
                          zref.astype(f32), visible, 3.0, 14.0, int(seed))


def depth_error_rate(depth_est: np.ndarray, scene: MultiViewScene, rel: float = 0.01) -> float:
    """Percentage of visible reference pixels with |z - z_gt| > rel * z_gt."""
    m = scene.visible
    err = np.abs(np.asarray(depth_est, np.float64)[m] - scene.depth[m])
    err[~np.isfinite(err)] = np.inf
    return float(100.0 * np.mean(err > rel * scene.depth[m]))


def multiview_1242x375(seed: int = 7) -> MultiViewScene:
    """Reference view and two source views of 1242x375, 4 rectangles."""
    return make_multiview(1242, 375, 2, seed)
