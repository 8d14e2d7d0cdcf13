"""Torch-tensor wrappers over the C-ABI, one per entry point of include/pmh_capi.h.

Tensors are the device memory: every argument must be a contiguous CUDA
(HIP) tensor of the stated dtype; results are allocated with torch.  Work is
enqueued on torch's current stream through a context bound to that stream,
so ordering with surrounding torch work is the stream order.
"""
from __future__ import annotations

import ctypes
import threading

import torch

from . import _lib

_ctx_cache: dict = {}
_ctx_lock = threading.Lock()


def context() -> _lib.Context:
    s = torch.cuda.current_stream()
    key = (s.device.index, s.cuda_stream, threading.get_ident())
    with _ctx_lock:
        c = _ctx_cache.get(key)
        if c is None:
            with torch.cuda.device(s.device):
                c = _lib.Context(s.cuda_stream)
            _ctx_cache[key] = c
    return c


def _ptr(t: torch.Tensor | None, dtype=torch.float32, name: str = "tensor"):
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError(f"{name} must be a device tensor")
    if t.dtype != dtype:
        raise ValueError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


def _hw(img: torch.Tensor):
    return int(img.shape[0]), int(img.shape[1])


def _fmt(img: torch.Tensor) -> int:
    """Image format of an input tensor: (H, W) gray or (H, W, 4) RGBA."""
    if img.dim() == 2:
        return _lib.FORMAT_GRAY
    if img.dim() == 3 and img.shape[2] == 4:
        return _lib.FORMAT_RGBA
    raise ValueError(f"input images are (H, W) gray or (H, W, 4) RGBA, got {tuple(img.shape)}")


def _call(fn, *args):
    _lib.check(fn(context().handle, *args))


def gradxy(img):
    """pmhstereo_gradxy (pmhstereo.cl:785-831): (H, W) -> (H, W, 4)."""
    H, W = _hw(img)
    out = torch.empty((H, W, 4), dtype=torch.float32, device=img.device)
    _call(_lib.load().pmh_gradxy, _ptr(img, name="img"), W, H, _ptr(out), _fmt(img))
    return out


def weight(img):
    """pmhstereo_weight (pmhstereo.cl:737-783): (H, W) -> (H, W)."""
    H, W = _hw(img)
    out = torch.empty((H, W), dtype=torch.float32, device=img.device)
    _call(_lib.load().pmh_weight, _ptr(img, name="img"), W, H, _ptr(out), _fmt(img))
    return out


def rng_draws(W, H, it, stride, count, device="cuda"):
    """MWC64X draws of RNG.cl as seeded at pmhstereo.cl:241 (stride 3) / :327 (45)."""
    out = torch.empty((H, W, count), dtype=torch.float32, device=device)
    _call(_lib.load().pmh_rng_draws, W, H, it, stride, count, _ptr(out))
    return out


def init(I1, G1, I2, G2, P, depthmax, lr):
    """pmhstereo_init (pmhstereo.cl:230-248)."""
    H, W = _hw(I1)
    out = torch.empty((H, W, 4), dtype=torch.float32, device=I1.device)
    _call(_lib.load().pmh_init, _ptr(I1, name="I1"), _ptr(G1, name="G1"), _ptr(I2, name="I2"),
          _ptr(G2, name="G2"), W, H, P, float(depthmax), int(lr), _ptr(out), _fmt(I1))
    return out


def cost(I1, G1, I2, G2, states, P):
    """patch_eval_color_grad_weighted (pmhstereo.cl:170-228) over a plane field."""
    H, W = _hw(I1)
    out = torch.empty((H, W), dtype=torch.float32, device=I1.device)
    _call(_lib.load().pmh_cost, _ptr(I1, name="I1"), _ptr(G1, name="G1"), _ptr(I2, name="I2"),
          _ptr(G2, name="G2"), W, H, P, _ptr(states, name="states"), _ptr(out), _fmt(I1))
    return out


def propagate(old, I1, G1, I2, G2, smooth, theta, P, depthmax, lr, it, viewin=None,
              viewsrc=None, draws=None, sweep=_lib.SWEEP_ALL, out=None):
    """pmhstereo_propagate_view (pmhstereo.cl:286-404) without the scatter."""
    H, W = _hw(I1)
    if out is None:
        out = old.clone()
    _call(_lib.load().pmh_propagate, _ptr(old, name="old"), _ptr(out, name="out"),
          _ptr(I1, name="I1"), _ptr(G1, name="G1"), _ptr(I2, name="I2"), _ptr(G2, name="G2"),
          _ptr(smooth, name="smooth"), float(theta), W, H, P, float(depthmax), int(lr), int(it),
          _ptr(viewin, torch.int32, "viewin"), _ptr(viewsrc, name="viewsrc"),
          _ptr(draws, name="draws"), int(sweep), _fmt(I1))
    return out


def viewbuf_build(states):
    """View scatter of pmhstereo.cl:391-401 -> (H, W, 5) int32 {n, src0..3}."""
    H, W = states.shape[:2]
    out = torch.zeros((H, W, 5), dtype=torch.int32, device=states.device)
    _call(_lib.load().pmh_viewbuf_build, _ptr(states, name="states"), W, H,
          _ptr(out, torch.int32))
    return out


def consistency(left, right, maxdispdiff, maxdeg, lr):
    """pmhstereo_consistency (pmhstereo.cl:428-464)."""
    H, W = left.shape[:2]
    out = torch.empty((H, W, 4), dtype=torch.float32, device=left.device)
    _call(_lib.load().pmh_consistency, _ptr(left, name="left"), _ptr(right, name="right"), W, H,
          float(maxdispdiff), float(maxdeg), int(lr), _ptr(out))
    return out


def fill_state(st, depthmax, lr):
    """pmhstereo_fill_state (pmhstereo.cl:466-510)."""
    H, W = st.shape[:2]
    out = torch.empty((H, W, 4), dtype=torch.float32, device=st.device)
    _call(_lib.load().pmh_fill_state, _ptr(st, name="in"), W, H, float(depthmax), int(lr),
          _ptr(out))
    return out


def pd_step(last, plast, f, w, lam):
    """One PDHuberWeighted step (PDHuberWeighted.cl:34-104); p is (H, W, 8)."""
    H, W = f.shape[:2]
    out = torch.empty((H, W, 4), dtype=torch.float32, device=f.device)
    pout = torch.empty((H, W, 8), dtype=torch.float32, device=f.device)
    _call(_lib.load().pmh_pd_step, _ptr(last, name="last"), _ptr(plast, name="plast"),
          _ptr(f, name="f"), _ptr(w, name="w"), W, H, float(lam), _ptr(out), _ptr(pout))
    return out, pout


def pdrof(u0, f, w, lam, iters):
    """PDROF::apply (PDROF.h:50-96); u0 is the warm start (not modified)."""
    H, W = f.shape[:2]
    u = u0.clone()
    _call(_lib.load().pmh_pdrof, _ptr(u, name="u"), _ptr(f, name="f"), _ptr(w, name="w"), W, H,
          float(lam), int(iters))
    return u


def occmap(left, right, maxdispdiff, lr, maxdeg=5.0):
    """pmhstereo_occmap (pmhstereo.cl:682-712) -> (H, W, 4) of (v, v, v, 1)."""
    H, W = left.shape[:2]
    out = torch.empty((H, W, 4), dtype=torch.float32, device=left.device)
    _call(_lib.load().pmh_occmap, _ptr(left, name="left"), _ptr(right, name="right"), W, H,
          float(maxdispdiff), float(maxdeg), int(lr), _ptr(out))
    return out


def normal_depth(st, depthmax, lr):
    """pmhstereo_normal_depth (pmhstereo.cl:663-680)."""
    H, W = st.shape[:2]
    out = torch.empty((H, W, 4), dtype=torch.float32, device=st.device)
    _call(_lib.load().pmh_normal_depth, _ptr(st, name="in"), W, H, float(depthmax), int(lr),
          _ptr(out))
    return out


def pd_step_inpaint(last, plast, f, w, mask, lam):
    """One PDHuberWeightedInpaint step (PDHuberWeightedInpaint.cl:34-104)."""
    H, W = f.shape[:2]
    out = torch.empty((H, W, 4), dtype=torch.float32, device=f.device)
    pout = torch.empty((H, W, 8), dtype=torch.float32, device=f.device)
    _call(_lib.load().pmh_pd_step_inpaint, _ptr(last, name="last"), _ptr(plast, name="plast"),
          _ptr(f, name="f"), _ptr(w, name="w"), _ptr(mask, name="mask"), W, H, float(lam),
          _ptr(out), _ptr(pout))
    return out, pout


def pdrof_inpaint(u0, f, w, mask, lam, iters):
    """PDROFInpaint::apply (PDROFInpaint.h:50-98); u0 is the warm start."""
    H, W = f.shape[:2]
    u = u0.clone()
    _call(_lib.load().pmh_pdrof_inpaint, _ptr(u, name="u"), _ptr(f, name="f"), _ptr(w, name="w"),
          _ptr(mask, name="mask"), W, H, float(lam), int(iters))
    return u


def depthmap_inpaint(left, right, patchsize, depthmax, iterations, viewsamples=4,
                     options: _lib.Options | None = None, states=False):
    """PMHuberStereo::depthMapInpaint (PMHuberStereo.cpp:293-507): disparity in px."""
    if left.shape != right.shape:
        raise ValueError("Left/Right stereo images inconsistent or incompatible memory type")
    if _fmt(left) == _lib.FORMAT_RGBA:
        return batch_depthmap_inpaint([(left, right, depthmax)], patchsize, iterations, options,
                                      states=states)[0]
    H, W = _hw(left)
    dmap = torch.empty((H, W), dtype=torch.float32, device=left.device)
    ls = torch.empty((H, W, 4), dtype=torch.float32, device=left.device) if states else None
    rs = torch.empty((H, W, 4), dtype=torch.float32, device=left.device) if states else None
    opts = options if options is not None else _lib.default_inpaint_options()
    _call(_lib.load().pmh_depthmap_inpaint, _ptr(left, name="left"), _ptr(right, name="right"),
          W, H, int(patchsize), float(depthmax), int(iterations), int(viewsamples), _ptr(dmap),
          ctypes.byref(opts), _ptr(ls), _ptr(rs))
    res = {"dmap": dmap}
    if states:
        res["left_states"], res["right_states"] = ls, rs
    return res


def fill_depthmap(st, scale):
    """pmhstereo_fill_depthmap (pmhstereo.cl:512-555)."""
    H, W = st.shape[:2]
    out = torch.empty((H, W), dtype=torch.float32, device=st.device)
    _call(_lib.load().pmh_fill_depthmap, _ptr(st, name="in"), W, H, float(scale), _ptr(out))
    return out


def fill_normalmap(st):
    """pmhstereo_fill_normalmap (pmhstereo.cl:557-603)."""
    H, W = st.shape[:2]
    out = torch.empty((H, W, 4), dtype=torch.float32, device=st.device)
    _call(_lib.load().pmh_fill_normalmap, _ptr(st, name="in"), W, H, _ptr(out))
    return out


def pyr_down(img):
    """2x2 box reduction of an input image (coarse-to-fine start), per channel."""
    H, W = _hw(img)
    out = torch.empty(((H + 1) // 2, (W + 1) // 2) + tuple(img.shape[2:]), dtype=torch.float32,
                      device=img.device)
    _call(_lib.load().pmh_pyr_down, _ptr(img, name="img"), W, H, _ptr(out), _fmt(img))
    return out


def plane_up(states, W, H):
    """Plane field of the half-resolution level carried to W x H."""
    h, w = states.shape[:2]
    out = torch.empty((H, W, 4), dtype=torch.float32, device=states.device)
    _call(_lib.load().pmh_plane_up, _ptr(states, name="in"), w, h, int(W), int(H), _ptr(out))
    return out


def cost_backward(I1, G1, I2, G2, states, gout, P, images=True):
    """Adjoint of the window cost: (dstates, dI1, dG1, dI2, dG2) for
    L = sum(gout * cost); image gradients are None unless ``images``.  dI1 and
    dI2 are shaped like the images (RGBA: the alpha channel gets none)."""
    H, W = _hw(I1)
    dev = I1.device
    dst = torch.empty((H, W, 4), dtype=torch.float32, device=dev)
    dI1 = torch.zeros(tuple(I1.shape), dtype=torch.float32, device=dev) if images else None
    dG1 = torch.zeros((H, W, 4), dtype=torch.float32, device=dev) if images else None
    dI2 = torch.zeros(tuple(I1.shape), dtype=torch.float32, device=dev) if images else None
    dG2 = torch.zeros((H, W, 4), dtype=torch.float32, device=dev) if images else None
    _call(_lib.load().pmh_cost_backward, _ptr(I1, name="I1"), _ptr(G1, name="G1"),
          _ptr(I2, name="I2"), _ptr(G2, name="G2"), W, H, P, _ptr(states, name="states"),
          _ptr(gout, name="gout"), _ptr(dst), _ptr(dI1), _ptr(dG1), _ptr(dI2), _ptr(dG2),
          _fmt(I1))
    return dst, dI1, dG1, dI2, dG2


def gradxy_backward(dG, dI=None):
    """Adjoint of pmhstereo_gradxy: returns dI + gradxy^T(dG)."""
    H, W = dG.shape[:2]
    out = torch.zeros((H, W), dtype=torch.float32, device=dG.device) if dI is None else dI.clone()
    _call(_lib.load().pmh_gradxy_backward, _ptr(dG, name="dG"), W, H, _ptr(out))
    return out


def grad_refine(states, I1, G1, I2, G2, smooth, theta, P, depthmax, lr, steps, radius=0.5):
    """Gradient-step plane refinement; returns the refined field."""
    H, W = _hw(I1)
    out = states.clone()
    _call(_lib.load().pmh_grad_refine, _ptr(out, name="states"), _ptr(I1, name="I1"),
          _ptr(G1, name="G1"), _ptr(I2, name="I2"), _ptr(G2, name="G2"),
          _ptr(smooth, name="smooth"), float(theta), W, H, P, float(depthmax), int(lr),
          int(steps), float(radius), _fmt(I1))
    return out


def depthmap(left, right, patchsize, depthmax, iterations, viewsamples=4, dscale=-1.0,
             options: _lib.Options | None = None, normals=False, states=False):
    """PMHuberStereo::depthMap (PMHuberStereo.cpp:64-291) on the device; RGBA
    (H, W, 4) inputs run as a batch of one pair."""
    if left.shape != right.shape:
        raise ValueError("Left/Right stereo images inconsistent or incompatible memory type")
    if _fmt(left) == _lib.FORMAT_RGBA:
        return batch_depthmap([(left, right, depthmax, dscale)], patchsize, iterations, options,
                              normals=normals, states=states)[0]
    H, W = _hw(left)
    dmap = torch.empty((H, W), dtype=torch.float32, device=left.device)
    nmap = torch.empty((H, W, 4), dtype=torch.float32, device=left.device) if normals else None
    ls = torch.empty((H, W, 4), dtype=torch.float32, device=left.device) if states else None
    rs = torch.empty((H, W, 4), dtype=torch.float32, device=left.device) if states else None
    opts = options if options is not None else _lib.default_options()
    _call(_lib.load().pmh_depthmap, _ptr(left, name="left"), _ptr(right, name="right"), W, H,
          int(patchsize), float(depthmax), int(iterations), int(viewsamples), float(dscale),
          _ptr(dmap), _ptr(nmap), ctypes.byref(opts), _ptr(ls), _ptr(rs))
    res = {"dmap": dmap}
    if normals:
        res["normals"] = nmap
    if states:
        res["left_states"], res["right_states"] = ls, rs
    return res


def batch_depthmap(pairs, patchsize, iterations, options: _lib.Options | None = None,
                   normals=False, states=False):
    """Many pairs in one run (pmh_batch_depthmap).  ``pairs`` is a list of
    (left, right, depthmax) or (left, right, depthmax, dscale); returns one
    result dict per pair, as ``depthmap`` does."""
    opts = options if options is not None else _lib.default_options()
    return _batch(_lib.load().pmh_batch_depthmap, pairs, patchsize, iterations, opts, normals,
                  states)


def _batch(fn, pairs, patchsize, iterations, opts, normals, states):
    if not pairs:
        raise ValueError("no pairs")
    arr = (_lib.Pair * len(pairs))()
    results = []
    for k, pr in enumerate(pairs):
        left, right, dmax = pr[0], pr[1], float(pr[2])
        dscale = float(pr[3]) if len(pr) > 3 else -1.0
        if left.shape != right.shape:
            raise ValueError(f"pair {k}: left/right images inconsistent")
        H, W = _hw(left)
        dev = left.device
        res = {"dmap": torch.empty((H, W), dtype=torch.float32, device=dev)}
        if normals:
            res["normals"] = torch.empty((H, W, 4), dtype=torch.float32, device=dev)
        if states:
            res["left_states"] = torch.empty((H, W, 4), dtype=torch.float32, device=dev)
            res["right_states"] = torch.empty((H, W, 4), dtype=torch.float32, device=dev)
        q = arr[k]
        q.left = _ptr(left, name="left").value
        q.right = _ptr(right, name="right").value
        q.width, q.height, q.depthmax, q.dscale = W, H, dmax, dscale
        q.format = _fmt(left)
        q.dmap = res["dmap"].data_ptr()
        q.normalmap = res["normals"].data_ptr() if normals else None
        q.left_states = res["left_states"].data_ptr() if states else None
        q.right_states = res["right_states"].data_ptr() if states else None
        results.append(res)
    _call(fn, arr, len(pairs), int(patchsize), int(iterations), ctypes.byref(opts))
    return results

def batch_depthmap_inpaint(pairs, patchsize, iterations, options: _lib.Options | None = None,
                           states=False):
    """depthMapInpaint over many pairs in one run; ``pairs`` as for batch_depthmap."""
    opts = options if options is not None else _lib.default_inpaint_options()
    return _batch(_lib.load().pmh_batch_depthmap_inpaint, pairs, patchsize, iterations, opts,
                  False, states)


# ------------------------------------------------------------- multi-view ----
class MVScene:
    """A prepared multi-view scene (pmh_mv_scene_create) on torch's current
    device: the reference gray image ``ref`` (H, W) with camera (K, T) and 1-4
    source views ``srcs`` with cameras (Ks[v], Ts[v]); K 3x3, T = [R | t] 3x4
    world -> camera.  Conventions of the particle fields (H, W, 4) = (nx, ny,
    d, cost): include/pmh_capi.h.  The methods wrap the pmh_mv_* entry points
    (the draft kernels of pmmvstereo.cl:97-304 and Plane.cl:60-86)."""

    def __init__(self, ref, K, T, srcs, Ks, Ts):
        import numpy as np
        if ref.dim() != 2:
            raise ValueError("multi-view images are (H, W) gray")
        self.H, self.W = _hw(ref)
        self.nsrc = len(srcs)
        self.device = ref.device

        def view(img, k, t):
            if tuple(img.shape) != (self.H, self.W):
                raise ValueError("all views must have the reference view's size")
            v = _lib.MVView()
            v.image = _ptr(img, name="image").value
            v.K[:] = [float(a) for a in np.asarray(k, np.float32).reshape(9)]
            v.T[:] = [float(a) for a in np.asarray(t, np.float32).reshape(12)]
            return v

        r = view(ref, K, T)
        arr = (_lib.MVView * max(self.nsrc, 1))()
        for i, (img, k, t) in enumerate(zip(srcs, Ks, Ts)):
            arr[i] = view(img, k, t)
        self._ctx = context()
        h = ctypes.c_void_p()
        _lib.check(_lib.load().pmh_mv_scene_create(self._ctx.handle, ctypes.byref(r), arr,
                                                   self.nsrc, self.W, self.H, ctypes.byref(h)))
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().pmh_mv_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _field(self):
        return torch.empty((self.H, self.W, 4), dtype=torch.float32, device=self.device)

    def init(self, P, dmin, dmax):
        """pmmv3_init (pmmvstereo.cl:170-195)."""
        out = self._field()
        _call(_lib.load().pmh_mv_init, self.handle, int(P), float(dmin), float(dmax), _ptr(out))
        return out

    def cost(self, particles, P):
        """pmmv3_patch_eval (pmmvstereo.cl:97-167) over a particle field."""
        out = torch.empty((self.H, self.W), dtype=torch.float32, device=self.device)
        _call(_lib.load().pmh_mv_cost, self.handle, int(P), _ptr(particles, name="particles"),
              _ptr(out))
        return out

    def propagate(self, old, P, dmin, dmax, it, draws=None, sweep=_lib.SWEEP_ALL, out=None,
                  refine=2):
        """pmmv3_iteration (pmmvstereo.cl:197-304)."""
        if out is None:
            out = old.clone()
        _call(_lib.load().pmh_mv_propagate, self.handle, _ptr(old, name="old"),
              _ptr(out, name="out"), int(P), float(dmin), float(dmax), int(it),
              _ptr(draws, name="draws"), int(sweep), int(refine))
        return out

    def homography(self, particles):
        """planef_homography (Plane.cl:60-86) of each particle: (n, nsrc, 3, 3)."""
        flat = particles.reshape(-1, 4).contiguous()
        out = torch.empty((flat.shape[0], self.nsrc, 3, 3), dtype=torch.float32,
                          device=self.device)
        _call(_lib.load().pmh_mv_homography, self.handle, _ptr(flat, name="particles"),
              int(flat.shape[0]), _ptr(out))
        return out

    def depth(self, particles):
        out = torch.empty((self.H, self.W), dtype=torch.float32, device=self.device)
        _call(_lib.load().pmh_mv_depth, self.handle, _ptr(particles, name="particles"),
              _ptr(out))
        return out

    def depthmap(self, P, dmin, dmax, iterations, sweep=_lib.SWEEP_REDBLACK, refine=2,
                 particles=False):
        """init + passes + depth (pmh_mv_depthmap) -> depth (H, W), and the
        final particle field if ``particles``."""
        depth = torch.empty((self.H, self.W), dtype=torch.float32, device=self.device)
        parts = self._field() if particles else None
        _call(_lib.load().pmh_mv_depthmap, self.handle, int(P), float(dmin), float(dmax),
              int(iterations), int(sweep), int(refine), _ptr(depth), _ptr(parts))
        return (depth, parts) if particles else depth
