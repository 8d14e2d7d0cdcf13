"""pmh_pdrof_noalloc (PDROF::apply with its scratch in a file-scope device
array) against pmh_pdrof through the C-ABI: bit for bit, on two calls in a row
(the second smaller, so it runs on the first call's leftover scratch), odd and
even step counts; sizes beyond the array are refused before any work."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(fn, *args):
    from cvt_amd import _lib
    _lib.check(fn(*args))


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def test_pdrof_noalloc_matches_pdrof():
    import torch
    from cvt_amd import _lib
    lib = _lib.load()
    ctx = _lib.Context(torch.cuda.current_stream().cuda_stream)
    rng = np.random.default_rng(21)
    for W, H, iters, lam in ((300, 200, 250, 5.0), (97, 45, 7, 17.9)):
        print(f"pdrof_noalloc {W}x{H} iters={iters}", flush=True)
        f = rng.random((H, W, 4)).astype(np.float32)
        f[..., 3] = 1
        w = (0.05 + rng.random((H, W))).astype(np.float32)
        u0 = rng.random((H, W, 4)).astype(np.float32)
        ft, wt = torch.from_numpy(f).cuda(), torch.from_numpy(w).cuda()
        ref = torch.from_numpy(u0).cuda()
        got = torch.from_numpy(u0).cuda()
        torch.cuda.synchronize()
        _call(lib.pmh_pdrof, ctx.handle, _ptr(ref), _ptr(ft), _ptr(wt), W, H, ctypes.c_float(lam),
              iters)
        _call(lib.pmh_pdrof_noalloc, ctx.handle, _ptr(got), _ptr(ft), _ptr(wt), W, H,
              ctypes.c_float(lam), iters)
        ctx.synchronize()
        assert torch.equal(got, ref), (W, H, iters)
        assert not torch.equal(got, torch.from_numpy(u0).cuda())  # it did smooth
    # beyond the scratch: refused with a clear error, nothing enqueued
    big = torch.zeros((1025, 1024, 4), device="cuda")
    bw = torch.zeros((1025, 1024), device="cuda")
    st = lib.pmh_pdrof_noalloc(ctx.handle, _ptr(big), _ptr(big), _ptr(bw), 1024, 1025,
                               ctypes.c_float(5.0), 4)
    assert st == -1  # PMH_ERR_INVALID
    assert b"exceeds the no-alloc scratch" in lib.pmh_last_error()
    ctx.close()
