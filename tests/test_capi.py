"""The C-ABI library loads and exports every symbol include/pmh_capi.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "pmh_capi.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(pmh_[a-z0-9_]+)\s*\(", src)))


def test_header_and_binding_agree():
    from cvt_amd import _lib
    assert sorted(_lib.EXPORTS) == declared_symbols()


def test_library_exports_all_declared():
    from cvt_amd import _lib
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_default_options_are_reference_constants():
    from cvt_amd import _lib
    o = _lib.default_options()
    assert o.maxdispdiff == 0.5 and o.maxanglediff == 10.0 and o.pd_iters == 250
    assert o.sweep == _lib.SWEEP_ALL and o.grad_steps == 0


def test_null_context_is_an_error_not_a_crash():
    from cvt_amd import _lib
    L = _lib.load()
    st = L.pmh_gradxy(None, None, 4, 4, None)
    assert st == -1
    assert b"null pmh_ctx" in L.pmh_last_error()
    with pytest.raises(_lib.PmhError):
        _lib.check(st)


def test_missing_library_fails_loudly(tmp_path):
    from cvt_amd import _lib
    with pytest.raises(_lib.PmhError):
        ctypes_path = str(tmp_path / "nope.so")
        saved = _lib._lib
        _lib._lib = None
        try:
            _lib.load(ctypes_path)
        finally:
            _lib._lib = saved
