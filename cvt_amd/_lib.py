"""ctypes binding of libpmh_amd.so (declarations in include/pmh_capi.h).

The HIP library is the only implementation: if it is missing or cannot be
loaded, every call raises.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpmh_amd.so")

PMH_OK = 0
SWEEP_ALL, SWEEP_RED, SWEEP_BLACK, SWEEP_REDBLACK = 0, 1, 2, 3
FORMAT_GRAY, FORMAT_RGBA = 0, 1

# every function include/pmh_capi.h declares
EXPORTS = (
    "pmh_create", "pmh_destroy", "pmh_synchronize", "pmh_last_error", "pmh_default_options",
    "pmh_depthmap", "pmh_gradxy", "pmh_weight", "pmh_rng_draws", "pmh_init", "pmh_cost",
    "pmh_propagate", "pmh_viewbuf_build", "pmh_consistency", "pmh_fill_state", "pmh_pd_step",
    "pmh_pdrof", "pmh_fill_depthmap", "pmh_fill_normalmap", "pmh_cost_backward",
    "pmh_gradxy_backward", "pmh_grad_refine", "pmh_profile_enable", "pmh_profile_reset",
    "pmh_profile_read", "pmh_profile_only", "pmh_profile_candidates", "pmh_batch_depthmap",
    "pmh_autotune_enable", "pmh_autotune_sample", "pmh_tuning_load", "pmh_tuning_save",
    "pmh_graph_stats", "pmh_pyr_down", "pmh_plane_up",
    "pmh_depthmap_inpaint", "pmh_default_inpaint_options", "pmh_batch_depthmap_inpaint",
    "pmh_occmap", "pmh_normal_depth", "pmh_pd_step_inpaint", "pmh_pdrof_inpaint",
    "pmh_pdrof_noalloc", "pmh_pfm_save", "pmh_cvtraw_save", "pmh_cvtraw_info",
    "pmh_cvtraw_load", "pmh_io_last_error",
    "pmh_mv_scene_create", "pmh_mv_scene_destroy", "pmh_mv_init", "pmh_mv_cost",
    "pmh_mv_propagate", "pmh_mv_homography", "pmh_mv_depth", "pmh_mv_depthmap",
)


class PmhError(RuntimeError):
    pass


class Options(ctypes.Structure):
    _fields_ = [("maxdispdiff", ctypes.c_float), ("maxanglediff", ctypes.c_float),
                ("pd_iters", ctypes.c_int), ("sweep", ctypes.c_int),
                ("grad_steps", ctypes.c_int), ("grad_step", ctypes.c_float),
                ("random_refine", ctypes.c_int), ("use_graph", ctypes.c_int),
                ("coarse_levels", ctypes.c_int), ("coarse_iters", ctypes.c_int),
                ("coarse_patch", ctypes.c_int), ("coarse_pd_iters", ctypes.c_int),
                ("coarse_random_refine", ctypes.c_int), ("random_neighbours", ctypes.c_int),
                ("coarse_random_neighbours", ctypes.c_int), ("coarse_joint_views", ctypes.c_int)]


class Pair(ctypes.Structure):
    """pmh_pair: one rectified pair of a batched run (device pointers)."""
    _fields_ = [("left", ctypes.c_void_p), ("right", ctypes.c_void_p),
                ("width", ctypes.c_int), ("height", ctypes.c_int), ("format", ctypes.c_int),
                ("depthmax", ctypes.c_float), ("dscale", ctypes.c_float),
                ("dmap", ctypes.c_void_p), ("normalmap", ctypes.c_void_p),
                ("left_states", ctypes.c_void_p), ("right_states", ctypes.c_void_p)]


class MVView(ctypes.Structure):
    """pmh_mv_view: a gray device image with its camera (K 3x3, T = [R | t]
    3x4 world -> camera, row-major)."""
    _fields_ = [("image", ctypes.c_void_p), ("K", ctypes.c_float * 9),
                ("T", ctypes.c_float * 12)]


_lib = None
_lock = threading.Lock()


def load(path: str = LIB_PATH):
    """Load the HIP library (raises if it is absent: no fallback path exists)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(path):
                raise PmhError(f"{path} is missing: build it with __graft_entry__.build() "
                               "(make -C cvt_amd/csrc)")
            L = ctypes.CDLL(path)
            _declare(L)
            _lib = L
    return _lib


def _declare(L):
    i, f, vp = ctypes.c_int, ctypes.c_float, ctypes.c_void_p
    pp = ctypes.POINTER(ctypes.c_void_p)
    sig = {
        "pmh_create": [pp, vp],
        "pmh_destroy": [vp],
        "pmh_synchronize": [vp],
        "pmh_depthmap": [vp, vp, vp, i, i, i, f, i, i, f, vp, vp, ctypes.POINTER(Options), vp, vp],
        "pmh_gradxy": [vp, vp, i, i, vp, i],
        "pmh_weight": [vp, vp, i, i, vp, i],
        "pmh_rng_draws": [vp, i, i, i, i, i, vp],
        "pmh_init": [vp, vp, vp, vp, vp, i, i, i, f, i, vp, i],
        "pmh_cost": [vp, vp, vp, vp, vp, i, i, i, vp, vp, i],
        "pmh_propagate": [vp, vp, vp, vp, vp, vp, vp, vp, f, i, i, i, f, i, i, vp, vp, vp, i, i],
        "pmh_viewbuf_build": [vp, vp, i, i, vp],
        "pmh_consistency": [vp, vp, vp, i, i, f, f, i, vp],
        "pmh_fill_state": [vp, vp, i, i, f, i, vp],
        "pmh_pd_step": [vp, vp, vp, vp, vp, i, i, f, vp, vp],
        "pmh_pdrof": [vp, vp, vp, vp, i, i, f, i],
        "pmh_pdrof_noalloc": [vp, vp, vp, vp, i, i, f, i],
        "pmh_fill_depthmap": [vp, vp, i, i, f, vp],
        "pmh_fill_normalmap": [vp, vp, i, i, vp],
        "pmh_cost_backward": [vp, vp, vp, vp, vp, i, i, i, vp, vp, vp, vp, vp, vp, vp, i],
        "pmh_gradxy_backward": [vp, vp, i, i, vp],
        "pmh_grad_refine": [vp, vp, vp, vp, vp, vp, vp, f, i, i, i, f, i, i, f, i],
        "pmh_batch_depthmap": [vp, ctypes.POINTER(Pair), i, i, i, ctypes.POINTER(Options)],
        "pmh_autotune_enable": [vp, i],
        "pmh_profile_only": [vp, ctypes.c_char_p],
        "pmh_profile_candidates": [vp, ctypes.POINTER(ctypes.c_longlong)],
        "pmh_autotune_sample": [vp, i],
        "pmh_tuning_load": [vp, ctypes.c_char_p],
        "pmh_tuning_save": [vp, ctypes.c_char_p],
        "pmh_graph_stats": [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_longlong)],
        "pmh_pyr_down": [vp, vp, i, i, vp, i],
        "pmh_depthmap_inpaint": [vp, vp, vp, i, i, i, f, i, i, vp, ctypes.POINTER(Options), vp,
                                 vp],
        "pmh_batch_depthmap_inpaint": [vp, ctypes.POINTER(Pair), i, i, i,
                                       ctypes.POINTER(Options)],
        "pmh_occmap": [vp, vp, vp, i, i, f, f, i, vp],
        "pmh_normal_depth": [vp, vp, i, i, f, i, vp],
        "pmh_pd_step_inpaint": [vp, vp, vp, vp, vp, vp, i, i, f, vp, vp],
        "pmh_pdrof_inpaint": [vp, vp, vp, vp, vp, i, i, f, i],
        "pmh_plane_up": [vp, vp, i, i, i, i, vp],
        "pmh_mv_scene_create": [vp, ctypes.POINTER(MVView), ctypes.POINTER(MVView), i, i, i,
                                pp],
        "pmh_mv_scene_destroy": [vp],
        "pmh_mv_init": [vp, vp, i, f, f, vp],
        "pmh_mv_cost": [vp, vp, i, vp, vp],
        "pmh_mv_propagate": [vp, vp, vp, vp, i, f, f, i, vp, i, i],
        "pmh_mv_homography": [vp, vp, vp, i, vp],
        "pmh_mv_depth": [vp, vp, vp, vp],
        "pmh_mv_depthmap": [vp, vp, i, f, f, i, i, i, vp, vp],
        "pmh_profile_enable": [vp, i],
        "pmh_profile_reset": [vp],
        "pmh_profile_read": [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                             ctypes.POINTER(ctypes.c_longlong)],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    L.pmh_last_error.argtypes = []
    L.pmh_last_error.restype = ctypes.c_char_p
    L.pmh_io_last_error.argtypes = []
    L.pmh_io_last_error.restype = ctypes.c_char_p
    sz, cp = ctypes.c_size_t, ctypes.c_char_p
    L.pmh_pfm_save.argtypes = [cp, vp, i, i, i]
    L.pmh_cvtraw_save.argtypes = [cp, vp, i, i, sz, i]
    L.pmh_cvtraw_info.argtypes = [cp, ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(sz),
                                  ctypes.POINTER(i)]
    L.pmh_cvtraw_load.argtypes = [cp, vp, sz]
    for name in ("pmh_pfm_save", "pmh_cvtraw_save", "pmh_cvtraw_info", "pmh_cvtraw_load"):
        getattr(L, name).restype = ctypes.c_int
    L.pmh_default_options.argtypes = [ctypes.POINTER(Options)]
    L.pmh_default_options.restype = None
    L.pmh_default_inpaint_options.argtypes = [ctypes.POINTER(Options)]
    L.pmh_default_inpaint_options.restype = None


def check(status: int) -> None:
    if status != PMH_OK:
        msg = load().pmh_last_error().decode(errors="replace")
        raise PmhError(f"pmh status {status}: {msg}")


def default_inpaint_options(**overrides) -> Options:
    """Options of depthMapInpaint with the reference's constants (occlusion map
    0.5 px / 5 degrees, 100 smoother steps)."""
    o = Options()
    load().pmh_default_inpaint_options(ctypes.byref(o))
    for k, v in overrides.items():
        setattr(o, k, v)
    return o


def default_options(**overrides) -> Options:
    o = Options()
    load().pmh_default_options(ctypes.byref(o))
    for k, v in overrides.items():
        setattr(o, k, v)
    return o


class Context:
    """A pmh_ctx bound to one HIP stream (0 / None: the default stream)."""

    def __init__(self, stream: int | None = None):
        self._lib = load()
        h = ctypes.c_void_p()
        check(self._lib.pmh_create(ctypes.byref(h), ctypes.c_void_p(stream or 0)))
        self.handle = h

    def close(self):
        if self.handle:
            self._lib.pmh_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        check(self._lib.pmh_synchronize(self.handle))

    def profile(self, enable: bool = True):
        check(self._lib.pmh_profile_enable(self.handle, 1 if enable else 0))

    def profile_reset(self):
        check(self._lib.pmh_profile_reset(self.handle))

    def autotune(self, enable: bool = True):
        check(self._lib.pmh_autotune_enable(self.handle, 1 if enable else 0))

    def autotune_sample(self, tiles: int = 0):
        """Largest work-list sample the autotuner times (0: the defaults)."""
        check(self._lib.pmh_autotune_sample(self.handle, int(tiles)))

    def tuning_load(self, path: str):
        check(self._lib.pmh_tuning_load(self.handle, os.fsencode(path)))

    def tuning_save(self, path: str):
        check(self._lib.pmh_tuning_save(self.handle, os.fsencode(path)))

    def graph_stats(self):
        """(graphs cached, graph replays so far)."""
        g, n = ctypes.c_int(), ctypes.c_longlong()
        check(self._lib.pmh_graph_stats(self.handle, ctypes.byref(g), ctypes.byref(n)))
        return g.value, n.value

    def profile_only(self, kernel: str | None = None):
        """Bracket only the launches of this kernel name (None: all)."""
        check(self._lib.pmh_profile_only(self.handle, kernel.encode() if kernel else None))

    def profile_candidates(self) -> int:
        """Window-cost candidates of the algorithm in the profiled propagate launches."""
        n = ctypes.c_longlong()
        check(self._lib.pmh_profile_candidates(self.handle, ctypes.byref(n)))
        return n.value

    def profile_read(self, kernel: str):
        """(total milliseconds, launches) of a kernel since the last reset."""
        ms, n = ctypes.c_double(), ctypes.c_longlong()
        check(self._lib.pmh_profile_read(self.handle, kernel.encode(), ctypes.byref(ms),
                                         ctypes.byref(n)))
        return ms.value, n.value
