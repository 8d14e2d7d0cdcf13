"""Python mirror of cvt::PMHuberStereo (cvt/vision/PMHuberStereo.h:35-59).

``depthMap`` keeps the reference's argument list and meaning.  Python has no
output references, so the disparity map (and the optional normal map) are
returned.  Inputs may be numpy arrays or device tensors; mismatched sizes raise
as the reference throws at PMHuberStereo.cpp:65-68.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib, ops


class PMHuberStereo:
    def __init__(self, options: _lib.Options | None = None):
        _lib.load()  # fails loudly when the HIP library is missing
        self.options = options if options is not None else _lib.default_options()

    @staticmethod
    def _device(img) -> torch.Tensor:
        if isinstance(img, torch.Tensor):
            t = img
        else:
            t = torch.from_numpy(np.ascontiguousarray(img, dtype=np.float32))
        if not (t.dim() == 2 or (t.dim() == 3 and t.shape[2] == 4)):
            raise ValueError("PMHuberStereo: images must be GRAY_FLOAT (H, W) or RGBA_FLOAT "
                             "(H, W, 4)")
        return t.to(device="cuda", dtype=torch.float32).contiguous()

    def depthMap(self, left, right, patchsize: int, depthmax: float, iterations: int,
                 viewsamples: int = 4, dscale: float = -1.0, normalmap: bool = False):
        """PMHuberStereo::depthMap (PMHuberStereo.cpp:64-291).

        Returns the disparity map |x' - x| * dscale (dscale <= 0: 1/depthmax), or
        (dmap, normalmap) when ``normalmap`` is true."""
        L, R = self._device(left), self._device(right)
        if L.shape != R.shape:
            raise ValueError("Left/Right stereo images inconsistent or incompatible memory type")
        res = ops.depthmap(L, R, patchsize, depthmax, iterations, viewsamples, dscale,
                           self.options, normals=normalmap)
        if normalmap:
            return res["dmap"], res["normals"]
        return res["dmap"]

    def _inpaint_options(self) -> _lib.Options:
        """The inpainting method's options: the reference's inpainting
        constants (occlusion tolerances, smoother steps) and every other field
        from this object's options."""
        opts = _lib.default_inpaint_options()
        for name, _ in _lib.Options._fields_:
            if name not in ("maxdispdiff", "maxanglediff", "pd_iters"):
                setattr(opts, name, getattr(self.options, name))
        return opts

    def depthMapInpaint(self, left, right, patchsize: int, depthmax: float, iterations: int,
                        viewsamples: int = 4, options: _lib.Options | None = None):
        """PMHuberStereo::depthMapInpaint (PMHuberStereo.cpp:293-507): the
        disparity map in px (the reference fills it with scale 1, :500-504).
        ``options`` defaults to the reference's inpainting constants."""
        L, R = self._device(left), self._device(right)
        if L.shape != R.shape:
            raise ValueError("Left/Right stereo images inconsistent or incompatible memory type")
        opts = options if options is not None else self._inpaint_options()
        return ops.depthmap_inpaint(L, R, patchsize, depthmax, iterations, viewsamples,
                                    opts)["dmap"]
