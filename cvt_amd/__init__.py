"""MI355X (gfx950) slanted-plane PatchMatch stereo: the PMHuberStereo path of
mherb/cvt rebuilt as hand-written HIP kernels behind a C-ABI (libpmh_amd.so,
include/pmh_capi.h).

Modules: ``_lib`` (ctypes binding, no fallback), ``ops`` (one torch-tensor
wrapper per C entry point), ``stereo`` (the PMHuberStereo class), ``autograd``
(the window cost as an autograd pair), ``formats`` (PFM / CVTRaw files),
``synth`` (seeded synthetic pairs with ground truth, bad-pixel score).
"""
from ._lib import (FORMAT_GRAY, FORMAT_RGBA, Options, PmhError, SWEEP_ALL, SWEEP_BLACK,
                   SWEEP_RED, SWEEP_REDBLACK, default_inpaint_options, default_options, load)
from .stereo import PMHuberStereo

__all__ = ["FORMAT_GRAY", "FORMAT_RGBA", "Options", "PmhError", "PMHuberStereo", "SWEEP_ALL",
           "SWEEP_RED", "SWEEP_BLACK", "SWEEP_REDBLACK", "default_inpaint_options",
           "default_options", "load"]
