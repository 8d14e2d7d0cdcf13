"""ctypes binding of the CPU restatement in ``pmh_oracle.c``.

TEST INFRASTRUCTURE ONLY: the checker of the HIP path and the timed CPU
baseline.  Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
cpu_baseline leg import this package; the product path (``cvt_amd``) never
does.  Each wrapper names the reference function it restates (paths relative
to the mherb/cvt tree).

Parity pins for each function (tests/test_oracle_*.py):
  rng_*            -- two independent MWC64X implementations (double-and-add
                      of RNG.cl:49-59 vs 128-bit remainder), skip(n) == n steps,
                      closed-form first outputs.
  state_* algebra  -- closed forms: viewprop is an involution, transform of an
                      init plane at its own pixel is x - z, ref_normal_depth
                      recovers (n, z).
  gradxy / weight  -- brute-force FP64 stencils with a zero border.
  cost             -- FP64 brute force, integer-disparity fronto-parallel case
                      (bilinear exact), identical images -> 0, small-support
                      and non-finite planes -> 1e5.
  cost_grad_f64    -- central finite differences in FP64.
  propagate        -- brute-force fold over explicitly enumerated candidates.
  viewbuf_build    -- brute-force per-target enumeration.
  consistency/fill -- per-pixel brute force on hand-built fields.
  fill_normalmap   -- per-pixel scans restated from pmhstereo.cl:557-603.
  pd_step          -- constant field is a fixed point; FP64 per-pixel step;
                      with a constant weight the iteration converges to the
                      Huber-ROF minimiser found independently by scipy (pins
                      the projection grouping).
  pdrof            -- equals the PDROF.h:50-96 ping-pong over pd_step for odd
                      and even step counts.
  depthmap         -- equals a step-by-step composition of
                      PMHuberStereo.cpp:64-291 from the pinned kernels, bit for
                      bit, at 1, 2 and 8 iterations (theta > 0).
  mv_* (multi-view, pmh_oracle_mv.c; tests/test_oracle_mv.py)
    mv_homography  -- identity cameras -> I; fronto-parallel plane under a
                      translation -> the shift f t / d; FP64 projection of
                      back-projected plane points through random cameras.
    mv_bilinear    -- texels at integer positions, the FP64 textbook
                      interpolant elsewhere, linear ramps, clamped edges.
    mv_cost        -- 0 for a copied or integer-shifted source view under the
                      matching plane; 1e5 cases; FP64 brute force.
    mv_particle_*  -- closed forms; the init normal equals nd_state_init's.
    mv_propagate   -- brute-force enumeration of the draft's candidate order,
                      strict < fold; RNG streams == supplied draws.
    mv_depthmap    -- equals init + passes composed in the test; recovers a
                      rendered scene's depth.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_pmh.so")
_lib = None

F32P = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
F64P = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
U32P = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
I32P = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")

SWEEP_ALL, SWEEP_RED, SWEEP_BLACK = 0, 1, 2
PROP_DRAWS = 21


class _Opts(ctypes.Structure):
    _fields_ = [("maxdispdiff", ctypes.c_float), ("maxanglediff", ctypes.c_float),
                ("pd_iters", ctypes.c_int), ("sweep", ctypes.c_int),
                ("random_neighbours", ctypes.c_int), ("random_refine", ctypes.c_int)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    c_int, c_float, c_u64, c_u32 = ctypes.c_int, ctypes.c_float, ctypes.c_uint64, ctypes.c_uint32
    vp = ctypes.c_void_p
    L.orc_mwc_mulmod.restype = c_u64
    L.orc_mwc_mulmod.argtypes = [c_u64, c_u64]
    L.orc_mwc_powmod.restype = c_u64
    L.orc_mwc_powmod.argtypes = [c_u64, c_u64]
    L.orc_rng_init.argtypes = [c_u64, c_u32, U32P]
    L.orc_rng_init_u128.argtypes = [c_u64, c_u32, U32P]
    L.orc_rng_float.restype = c_float
    L.orc_rng_float.argtypes = [U32P]
    L.orc_rng_step.argtypes = [U32P]
    L.orc_rng_draws.argtypes = [c_u64, c_u32, c_int, F32P]
    L.orc_state_init.argtypes = [F32P, c_float, c_float, c_int, c_float, c_float, F32P]
    L.orc_state_viewprop.argtypes = [F32P, F32P]
    L.orc_state_transform.restype = c_float
    L.orc_state_transform.argtypes = [F32P, c_float, c_float]
    L.orc_state_to_ref_normal_depth.argtypes = [F32P, c_float, c_float, c_int, F32P]
    L.orc_state_refine.argtypes = [F32P, F32P, c_float, c_float, c_float, c_int, F32P]
    L.orc_state_to_normal.argtypes = [F32P, F32P]
    L.orc_gradxy.argtypes = [F32P, c_int, c_int, c_int, F32P]
    L.orc_weight.argtypes = [F32P, c_int, c_int, c_int, F32P]
    L.orc_factor_table.argtypes = [c_int, F32P]
    L.orc_cost.restype = c_float
    L.orc_cost.argtypes = [F32P, F32P, F32P, F32P, c_int, c_int, c_int, c_int, c_int, F32P,
                           c_int]
    L.orc_cost_field.argtypes = [F32P, F32P, F32P, F32P, c_int, c_int, c_int, F32P, c_int, F32P]
    L.orc_cost_grad_f64.restype = ctypes.c_double
    L.orc_cost_grad_f64.argtypes = [F32P, F32P, F32P, F32P, c_int, c_int, c_int, c_int, c_int,
                                    F64P, c_int, vp, vp, vp, vp, vp]
    L.orc_cost_grad_field_f64.argtypes = [F32P, F32P, F32P, F32P, c_int, c_int, c_int, F32P,
                                           c_int, F32P, F64P, F64P, F64P, F64P, F64P]
    L.orc_cost_f64.restype = ctypes.c_double
    L.orc_cost_f64.argtypes = [F32P, F32P, F32P, F32P, c_int, c_int, c_int, c_int, c_int, F64P,
                               c_int]
    L.orc_init.argtypes = [F32P, F32P, F32P, F32P, c_int, c_int, c_int, c_int, c_float, c_int,
                           F32P]
    L.orc_propagate.argtypes = [F32P, F32P, F32P, F32P, F32P, F32P, F32P, c_float, c_int, c_int,
                                c_int, c_int, c_float, c_int, c_int, vp, vp, vp, c_int]
    L.orc_propagate_ex.argtypes = L.orc_propagate.argtypes + [c_int, c_int]
    L.orc_viewbuf_build.argtypes = [F32P, c_int, c_int, I32P]
    L.orc_consistency.argtypes = [F32P, F32P, c_int, c_int, c_float, c_float, c_int, F32P]
    L.orc_fill_state.argtypes = [F32P, c_int, c_int, c_float, c_int, F32P]
    L.orc_pd_step.argtypes = [F32P, F32P, F32P, F32P, c_int, c_int, c_float, F32P, F32P]
    L.orc_pdrof.argtypes = [F32P, F32P, F32P, c_int, c_int, c_float, c_int]
    L.orc_fill_depthmap.argtypes = [F32P, c_int, c_int, c_float, F32P]
    L.orc_fill_normalmap.argtypes = [F32P, c_int, c_int, F32P]
    L.orc_grad_refine.argtypes = [F32P, F32P, F32P, F32P, F32P, F32P, c_int, c_int, c_int, c_int,
                                  c_float, c_int, c_float, c_int, c_float]
    L.orc_depthmap.argtypes = [F32P, F32P, c_int, c_int, c_int, c_int, c_float, c_int, c_float,
                               ctypes.POINTER(_Opts), F32P, vp, vp, vp]
    L.orc_occmap.argtypes = [F32P, F32P, c_int, c_int, c_float, c_float, c_int, F32P]
    L.orc_normal_depth.argtypes = [F32P, c_int, c_int, c_float, c_int, F32P]
    L.orc_pd_step_inpaint.argtypes = [F32P, F32P, F32P, F32P, F32P, c_int, c_int, c_float, F32P,
                                      F32P]
    L.orc_pdrof_inpaint.argtypes = [F32P, F32P, F32P, F32P, c_int, c_int, c_float, c_int]
    L.orc_depthmap_inpaint.argtypes = [F32P, F32P, c_int, c_int, c_int, c_int, c_float, c_int,
                                       ctypes.POINTER(_Opts), F32P, vp, vp]
    L.orc_mv_constants.argtypes = [F32P, F32P, c_int, F32P, F32P, F32P, F32P, F32P]
    L.orc_mv_homography.argtypes = [F32P, F32P, F32P, F32P, F32P]
    L.orc_mv_depth.restype = c_float
    L.orc_mv_depth.argtypes = [F32P, F32P, c_float, c_float]
    L.orc_mv_factor_table.argtypes = [c_int, F32P]
    L.orc_mv_cells.argtypes = [F32P, c_int, c_int, c_int, c_int, F32P]
    L.orc_mv_bilinear.restype = c_float
    L.orc_mv_bilinear.argtypes = [F32P, c_int, c_int, c_int, c_int, c_float, c_float]
    L.orc_mv_cost.restype = c_float
    L.orc_mv_cost.argtypes = [F32P, F32P, c_int, F32P, F32P, c_int, c_int, c_int, c_int, F32P,
                              c_int, F32P, F32P, F32P]
    L.orc_mv_particle_init.argtypes = [F32P, c_float, c_float, c_float, c_float, c_float, F32P,
                                       F32P]
    L.orc_mv_particle_refine.argtypes = [F32P, F32P, c_float, c_float, c_float, c_float, F32P,
                                         F32P]
    L.orc_mv_init.argtypes = [F32P, F32P, c_int, F32P, F32P, c_int, c_int, c_int, c_float,
                              c_float, F32P, F32P, F32P, F32P]
    L.orc_mv_propagate.argtypes = [F32P, F32P, F32P, F32P, c_int, F32P, F32P, c_int, c_int, c_int,
                                   c_float, c_float, c_int, F32P, F32P, F32P, vp, c_int, c_int]
    L.orc_mv_depthmap.argtypes = [F32P, F32P, c_int, F32P, F32P, c_int, c_int, c_int, c_float,
                                  c_float, c_int, c_int, c_int, F32P, F32P, F32P, F32P, vp]
    L.orc_num_threads.restype = c_int


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _hwc(img):
    """(H, W, nc) of an input image: (H, W) gray -> nc 1, (H, W, 4) RGBA -> nc 4."""
    if img.ndim == 2:
        return img.shape[0], img.shape[1], 1
    if img.ndim == 3 and img.shape[2] == 4:
        return img.shape[0], img.shape[1], 4
    raise ValueError(f"input images are (H, W) gray or (H, W, 4) RGBA, got {img.shape}")


def num_threads() -> int:
    return lib().orc_num_threads()


# ---- MWC64X (RNG.cl:35-156) ----
def mulmod(a: int, b: int) -> int:
    return int(lib().orc_mwc_mulmod(a, b))


def powmod(a: int, e: int) -> int:
    return int(lib().orc_mwc_powmod(a, e))


def rng_init(linpos: int, numdraws: int, fast: bool = False) -> np.ndarray:
    st = np.zeros(2, np.uint32)
    (lib().orc_rng_init_u128 if fast else lib().orc_rng_init)(linpos, numdraws, st)
    return st


def rng_float(st: np.ndarray) -> float:
    return float(lib().orc_rng_float(st))


def rng_step(st: np.ndarray) -> None:
    lib().orc_rng_step(st)


def rng_draws(linpos: int, numdraws: int, count: int) -> np.ndarray:
    out = np.zeros(count, np.float32)
    lib().orc_rng_draws(linpos, numdraws, count, out)
    return out


# ---- plane algebra (pmhstereo.cl:50-153) ----
def state_init(u, cx, cy, lr, normmul, depthmax):
    out = np.zeros(4, np.float32)
    lib().orc_state_init(_f32(u), cx, cy, lr, normmul, depthmax, out)
    return out


def state_viewprop(s):
    out = np.zeros(4, np.float32)
    lib().orc_state_viewprop(_f32(s), out)
    return out


def state_transform(s, cx, cy) -> float:
    return float(lib().orc_state_transform(_f32(s), cx, cy))


def state_to_ref_normal_depth(s, cx, cy, lr):
    out = np.zeros(4, np.float32)
    lib().orc_state_to_ref_normal_depth(_f32(s), cx, cy, lr, out)
    return out


def state_refine(u, s, cx, cy, depthmax, lr):
    out = np.zeros(4, np.float32)
    lib().orc_state_refine(_f32(u), _f32(s), cx, cy, depthmax, lr, out)
    return out


def state_to_normal(s):
    out = np.zeros(3, np.float32)
    lib().orc_state_to_normal(_f32(s), out)
    return out


# ---- images ----
def gradxy(img):
    img = _f32(img)
    H, W, nc = _hwc(img)
    out = np.zeros((H, W, 4), np.float32)
    lib().orc_gradxy(img, W, H, nc, out)
    return out


def weight(img):
    img = _f32(img)
    H, W, nc = _hwc(img)
    out = np.zeros((H, W), np.float32)
    lib().orc_weight(img, W, H, nc, out)
    return out


def factor_table(P):
    out = np.zeros((2 * P + 1, 2 * P + 1), np.float32)
    lib().orc_factor_table(P, out)
    return out


def cost(I1, G1, I2, G2, x, y, s, P) -> float:
    H, W, nc = _hwc(I1)
    return float(lib().orc_cost(_f32(I1), _f32(G1), _f32(I2), _f32(G2), W, H, nc, x, y,
                                _f32(s), P))


def cost_field(I1, G1, I2, G2, states, P):
    H, W, nc = _hwc(I1)
    out = np.zeros((H, W), np.float32)
    lib().orc_cost_field(_f32(I1), _f32(G1), _f32(I2), _f32(G2), W, H, nc, _f32(states), P, out)
    return out


def cost_f64(I1, G1, I2, G2, x, y, s, P) -> float:
    """FP64 window cost; gray (H, W) or RGBA (H, W, 4) images."""
    H, W, nc = _hwc(I1)
    return float(lib().orc_cost_f64(_f32(I1), _f32(G1), _f32(I2), _f32(G2), W, H, nc, x, y,
                                    np.ascontiguousarray(s, np.float64), P))


def _img_shape(H, W, nc):
    return (H, W) if nc == 1 else (H, W, nc)


def cost_grad_f64(I1, G1, I2, G2, x, y, s, P, images=True):
    """FP64 cost of one pixel and its gradient w.r.t. the plane and images
    (dI1, dI2 shaped like the images; the alpha channel of RGBA gets none)."""
    H, W, nc = _hwc(I1)
    dplane = np.zeros(3, np.float64)
    dI1 = np.zeros(_img_shape(H, W, nc), np.float64) if images else None
    dG1 = np.zeros((H, W, 4), np.float64) if images else None
    dI2 = np.zeros(_img_shape(H, W, nc), np.float64) if images else None
    dG2 = np.zeros((H, W, 4), np.float64) if images else None
    c = lib().orc_cost_grad_f64(_f32(I1), _f32(G1), _f32(I2), _f32(G2), W, H, nc, x, y,
                                np.ascontiguousarray(s, np.float64), P, _ptr(dplane),
                                _ptr(dI1), _ptr(dG1), _ptr(dI2), _ptr(dG2))
    return float(c), dplane, dI1, dG1, dI2, dG2


def cost_grad_field_f64(I1, G1, I2, G2, states, P, gout):
    """FP64 gradient of sum(gout * cost) w.r.t. every plane (H, W, 3) and the images."""
    H, W, nc = _hwc(I1)
    dp = np.zeros((H, W, 3), np.float64)
    dI1 = np.zeros(_img_shape(H, W, nc), np.float64)
    dG1 = np.zeros((H, W, 4), np.float64)
    dI2 = np.zeros(_img_shape(H, W, nc), np.float64)
    dG2 = np.zeros((H, W, 4), np.float64)
    lib().orc_cost_grad_field_f64(_f32(I1), _f32(G1), _f32(I2), _f32(G2), W, H, nc, _f32(states),
                                  P, _f32(gout), dp, dI1, dG1, dI2, dG2)
    return dp, dI1, dG1, dI2, dG2


def init(I1, G1, I2, G2, P, depthmax, lr):
    H, W, nc = _hwc(I1)
    out = np.zeros((H, W, 4), np.float32)
    lib().orc_init(_f32(I1), _f32(G1), _f32(I2), _f32(G2), W, H, nc, P, depthmax, lr, out)
    return out


def propagate(old, I1, G1, I2, G2, smooth, theta, P, depthmax, lr, it, viewin=None,
              viewsrc=None, draws=None, sweep=SWEEP_ALL, out=None, random_neighbours=6,
              random_refine=2):
    H, W, nc = _hwc(I1)
    old = _f32(old)
    if out is None:
        out = old.copy()
    vi = None if viewin is None else np.ascontiguousarray(viewin, np.int32)
    vs = None if viewsrc is None else _f32(viewsrc)
    dr = None if draws is None else _f32(draws)
    lib().orc_propagate_ex(old, out, _f32(I1), _f32(G1), _f32(I2), _f32(G2), _f32(smooth),
                           theta, W, H, nc, P, depthmax, lr, it, _ptr(vi), _ptr(vs), _ptr(dr),
                           sweep, random_neighbours, random_refine)
    return out


def grad_refine(states, I1, G1, I2, G2, smooth, theta, P, depthmax, lr, steps, radius=0.5):
    """The product's gradient plane refinement restated on the CPU (no
    reference counterpart): a refined copy of the plane field."""
    H, W, nc = _hwc(I1)
    out = _f32(states).copy()
    lib().orc_grad_refine(out, _f32(I1), _f32(G1), _f32(I2), _f32(G2), _f32(smooth), W, H, nc,
                          P, depthmax, lr, theta, steps, radius)
    return out


def viewbuf_build(states):
    H, W = states.shape[:2]
    out = np.zeros((H, W, 5), np.int32)
    lib().orc_viewbuf_build(_f32(states), W, H, out)
    return out


def consistency(left, right, maxdispdiff, maxdeg, lr):
    H, W = left.shape[:2]
    out = np.zeros((H, W, 4), np.float32)
    lib().orc_consistency(_f32(left), _f32(right), W, H, maxdispdiff, maxdeg, lr, out)
    return out


def fill_state(st, depthmax, lr):
    H, W = st.shape[:2]
    out = np.zeros((H, W, 4), np.float32)
    lib().orc_fill_state(_f32(st), W, H, depthmax, lr, out)
    return out


def pd_step(last, plast, f, w, lam):
    H, W = f.shape[:2]
    out = np.zeros((H, W, 4), np.float32)
    pout = np.zeros((H, W, 8), np.float32)
    lib().orc_pd_step(_f32(last), _f32(plast), _f32(f), _f32(w), W, H, lam, out, pout)
    return out, pout


def pdrof(u0, f, w, lam, iters):
    H, W = f.shape[:2]
    u = _f32(u0).copy()
    lib().orc_pdrof(u, _f32(f), _f32(w), W, H, lam, iters)
    return u


def fill_depthmap(st, scale):
    H, W = st.shape[:2]
    out = np.zeros((H, W), np.float32)
    lib().orc_fill_depthmap(_f32(st), W, H, scale, out)
    return out


def fill_normalmap(st):
    H, W = st.shape[:2]
    out = np.zeros((H, W, 4), np.float32)
    lib().orc_fill_normalmap(_f32(st), W, H, out)
    return out


def depthmap(left, right, P, depthmax, iterations, dscale=-1.0, maxdispdiff=0.5,
             maxanglediff=10.0, pd_iters=250, sweep=0, want_states=False, want_normals=False,
             random_neighbours=6, random_refine=2):
    """PMHuberStereo::depthMap (PMHuberStereo.cpp:64-291) on the CPU."""
    left, right = _f32(left), _f32(right)
    H, W, nc = _hwc(left)
    dmap = np.zeros((H, W), np.float32)
    nmap = np.zeros((H, W, 4), np.float32) if want_normals else None
    ls = np.zeros((H, W, 4), np.float32) if want_states else None
    rs = np.zeros((H, W, 4), np.float32) if want_states else None
    opts = _Opts(maxdispdiff, maxanglediff, pd_iters, sweep, random_neighbours, random_refine)
    lib().orc_depthmap(left, right, W, H, nc, P, depthmax, iterations, dscale,
                       ctypes.byref(opts), dmap, _ptr(nmap), _ptr(ls), _ptr(rs))
    res = {"dmap": dmap}
    if want_normals:
        res["normals"] = nmap
    if want_states:
        res["left_states"], res["right_states"] = ls, rs
    return res


# ---- depthMapInpaint (PMHuberStereo.cpp:293-507) and its kernels ----
def occmap(left, right, maxdiff, lr, maxdeg=5.0):
    """pmhstereo_occmap (pmhstereo.cl:682-712) -> (H, W, 4) of (v, v, v, 1)."""
    H, W = left.shape[:2]
    out = np.zeros((H, W, 4), np.float32)
    lib().orc_occmap(_f32(left), _f32(right), W, H, maxdiff, maxdeg, lr, out)
    return out


def normal_depth(st, depthmax, lr):
    """pmhstereo_normal_depth (pmhstereo.cl:663-680)."""
    H, W = st.shape[:2]
    out = np.zeros((H, W, 4), np.float32)
    lib().orc_normal_depth(_f32(st), W, H, depthmax, lr, out)
    return out


def pd_step_inpaint(last, plast, f, w, mask, lam):
    """One PDHuberWeightedInpaint step (PDHuberWeightedInpaint.cl:34-104)."""
    H, W = f.shape[:2]
    out = np.zeros((H, W, 4), np.float32)
    pout = np.zeros((H, W, 8), np.float32)
    lib().orc_pd_step_inpaint(_f32(last), _f32(plast), _f32(f), _f32(w), _f32(mask), W, H, lam,
                              out, pout)
    return out, pout


def pdrof_inpaint(u0, f, w, mask, lam, iters):
    """PDROFInpaint::apply (PDROFInpaint.h:50-98)."""
    H, W = f.shape[:2]
    u = _f32(u0).copy()
    lib().orc_pdrof_inpaint(u, _f32(f), _f32(w), _f32(mask), W, H, lam, iters)
    return u


def depthmap_inpaint(left, right, P, depthmax, iterations, occdiff=0.5, occdeg=5.0,
                     pd_iters=100, sweep=0, want_states=False):
    """PMHuberStereo::depthMapInpaint (PMHuberStereo.cpp:293-507) on the CPU."""
    left, right = _f32(left), _f32(right)
    H, W, nc = _hwc(left)
    dmap = np.zeros((H, W), np.float32)
    ls = np.zeros((H, W, 4), np.float32) if want_states else None
    rs = np.zeros((H, W, 4), np.float32) if want_states else None
    opts = _Opts(occdiff, occdeg, pd_iters, sweep, 6, 2)
    lib().orc_depthmap_inpaint(left, right, W, H, nc, P, depthmax, iterations,
                               ctypes.byref(opts), dmap, _ptr(ls), _ptr(rs))
    res = {"dmap": dmap}
    if want_states:
        res["left_states"], res["right_states"] = ls, rs
    return res


# ---- Multi-view PatchMatch (pmh_oracle_mv.c; draft pmmvstereo.cl:97-304,
# Plane.cl:60-86) ----
MV_DRAWS = 33


class MVViews:
    """Reference view (gray image, K, T) and 1-4 source views, with the
    gradient images and the view constants the multi-view functions take."""

    def __init__(self, ref, K, T, srcs, Ks, Ts):
        self.Ir = _f32(ref)
        H, W, nc = _hwc(self.Ir)
        if nc != 1:
            raise ValueError("multi-view inputs are gray (H, W) images")
        self.H, self.W = H, W
        self.nsrc = len(srcs)
        if not 1 <= self.nsrc <= 4:
            raise ValueError("1 to 4 source views")
        self.Is = _f32(np.stack([_f32(s) for s in srcs]))
        if self.Is.shape[1:] != (H, W):
            raise ValueError("source views must have the reference view's size")
        self.Gr = gradxy(self.Ir)
        self.Gs = _f32(np.stack([gradxy(s) for s in self.Is]))
        self.kinv, self.M, self.e = mv_constants(K, T, Ks, Ts)


def mv_constants(K, T, Ks, Ts):
    """(K_r^-1, M = K_s R K_r^-1 per view, e = K_s t per view)."""
    Ks, Ts = _f32(Ks), _f32(Ts)
    n = Ks.shape[0]
    kinv = np.zeros(9, np.float32)
    M = np.zeros((n, 9), np.float32)
    e = np.zeros((n, 3), np.float32)
    lib().orc_mv_constants(_f32(K), _f32(T), n, Ks, Ts, kinv, M, e)
    return kinv, M, e


def mv_homography(p, kinv, M, e):
    H = np.zeros(9, np.float32)
    lib().orc_mv_homography(_f32(p), _f32(kinv), _f32(M), _f32(e), H)
    return H.reshape(3, 3)


def mv_depth(p, kinv, x, y) -> float:
    return float(lib().orc_mv_depth(_f32(p), _f32(kinv), x, y))


def mv_cells(img, c=0):
    """Bilinear cells (t, dx, dy, dxy) of channel c of an (H, W) or (H, W, n) image."""
    img = _f32(img)
    H, W = img.shape[:2]
    nch = 1 if img.ndim == 2 else img.shape[2]
    out = np.zeros((H, W, 4), np.float32)
    lib().orc_mv_cells(img, W, H, nch, c, out)
    return out


def mv_bilinear(img, x, y, c=0) -> float:
    img = _f32(img)
    H, W = img.shape[:2]
    nch = 1 if img.ndim == 2 else img.shape[2]
    return float(lib().orc_mv_bilinear(img, W, H, nch, c, x, y))


def mv_factor_table(P):
    S = 2 * P + 1
    t = np.zeros(S * S, np.float32)
    lib().orc_mv_factor_table(P, t)
    return t.reshape(S, S)


def mv_cost(v: MVViews, x, y, p, P) -> float:
    return float(lib().orc_mv_cost(v.Ir, v.Gr, v.nsrc, v.Is, v.Gs, v.W, v.H, x, y, _f32(p), P,
                                   v.kinv, v.M, v.e))


def mv_particle_init(u, x, y, nm, dmin, dmax, kinv):
    out = np.zeros(4, np.float32)
    lib().orc_mv_particle_init(_f32(u), x, y, nm, dmin, dmax, _f32(kinv), out)
    return out


def mv_particle_refine(u, p, x, y, dmin, dmax, kinv):
    out = np.zeros(4, np.float32)
    lib().orc_mv_particle_refine(_f32(u), _f32(p), x, y, dmin, dmax, _f32(kinv), out)
    return out


def mv_init(v: MVViews, P, dmin, dmax):
    out = np.zeros((v.H, v.W, 4), np.float32)
    lib().orc_mv_init(v.Ir, v.Gr, v.nsrc, v.Is, v.Gs, v.W, v.H, P, dmin, dmax, v.kinv, v.M, v.e,
                      out)
    return out


def mv_propagate(v: MVViews, old, P, dmin, dmax, it, draws=None, sweep=SWEEP_ALL, out=None,
                 refine=2):
    old = _f32(old)
    if out is None:
        out = old.copy()
    dr = None if draws is None else _f32(draws)
    lib().orc_mv_propagate(old, out, v.Ir, v.Gr, v.nsrc, v.Is, v.Gs, v.W, v.H, P, dmin, dmax, it,
                           v.kinv, v.M, v.e, _ptr(dr), sweep, refine)
    return out


def mv_depthmap(v: MVViews, P, dmin, dmax, iterations, red_black=True, refine=2):
    depth = np.zeros((v.H, v.W), np.float32)
    parts = np.zeros((v.H, v.W, 4), np.float32)
    lib().orc_mv_depthmap(v.Ir, v.Gr, v.nsrc, v.Is, v.Gs, v.W, v.H, P, dmin, dmax, iterations,
                          1 if red_black else 0, refine, v.kinv, v.M, v.e, depth, _ptr(parts))
    return depth, parts
