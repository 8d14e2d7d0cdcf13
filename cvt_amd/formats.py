"""On-disk image formats of the path, through the library's C-ABI: the PFM
writer (plugins/isaver/PFM/PFMSaver.cpp:117-160) and the CVTRaw reader and
writer (plugins/iloader/CVTRaw/CVTRawLoader.cpp, plugins/isaver/CVTRaw/
CVTRawSaver.cpp).  Images are host arrays: (H, W) float32 gray or (H, W, 4)
float32 RGBA; device tensors are copied to the host first."""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib

# IFormat ids (cvt/gfx/IFormat.h:36-63) and bytes per pixel (IFormat.cpp:32-52)
GRAY_FLOAT, RGBA_FLOAT = 4, 12
_BPP = {1: 1, 2: 2, 3: 2, 4: 4, 5: 2, 6: 4, 7: 4, 8: 8, 9: 4, 10: 8, 11: 8, 12: 16, 13: 4,
        14: 8, 15: 8, 16: 16, 17: 1, 18: 1, 19: 1, 20: 2, 21: 2}


def _host(img) -> np.ndarray:
    if hasattr(img, "detach"):
        img = img.detach().cpu().numpy()
    a = np.ascontiguousarray(img, dtype=np.float32)
    if not (a.ndim == 2 or (a.ndim == 3 and a.shape[2] == 4)):
        raise ValueError(f"images are (H, W) gray or (H, W, 4) RGBA, got {a.shape}")
    return a


def _check(status: int) -> None:
    if status != _lib.PMH_OK:
        raise _lib.PmhError(_lib.load().pmh_io_last_error().decode(errors="replace"))


def save_pfm(path: str, img) -> None:
    """PFMSaver::save: gray -> "Pf", RGBA -> "PF" (alpha dropped)."""
    a = _host(img)
    fmt = _lib.FORMAT_GRAY if a.ndim == 2 else _lib.FORMAT_RGBA
    _check(_lib.load().pmh_pfm_save(os.fsencode(path), a.ctypes.data_as(ctypes.c_void_p),
                                    a.shape[1], a.shape[0], fmt))


def save_cvtraw(path: str, img) -> None:
    """CVTRawSaver::save of a GRAY_FLOAT or RGBA_FLOAT image."""
    a = _host(img)
    fid = GRAY_FLOAT if a.ndim == 2 else RGBA_FLOAT
    _check(_lib.load().pmh_cvtraw_save(os.fsencode(path), a.ctypes.data_as(ctypes.c_void_p),
                                       a.shape[1], a.shape[0], a.strides[0], fid))


def cvtraw_info(path: str):
    """(width, height, stride, IFormat id) of a CVTRaw file."""
    w, h, fid = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    st = ctypes.c_size_t()
    _check(_lib.load().pmh_cvtraw_info(os.fsencode(path), ctypes.byref(w), ctypes.byref(h),
                                       ctypes.byref(st), ctypes.byref(fid)))
    return w.value, h.value, st.value, fid.value


def load_cvtraw(path: str):
    """CVTRawLoader::load -> (array, IFormat id): (H, W) float32 for GRAY_FLOAT,
    (H, W, 4) float32 for RGBA_FLOAT, otherwise the raw rows as (H, W * bpp)
    uint8."""
    w, h, _, fid = cvtraw_info(path)
    if fid == GRAY_FLOAT:
        out = np.empty((h, w), np.float32)
    elif fid == RGBA_FLOAT:
        out = np.empty((h, w, 4), np.float32)
    else:
        out = np.empty((h, w * _BPP[fid]), np.uint8)
    _check(_lib.load().pmh_cvtraw_load(os.fsencode(path), out.ctypes.data_as(ctypes.c_void_p),
                                       out.strides[0]))
    return out, fid
