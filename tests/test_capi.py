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
    st = L.pmh_gradxy(None, None, 4, 4, None, 0)
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


def test_struct_layouts_match_the_header(tmp_path):
    """ctypes mirrors of pmh_options / pmh_pair against the C compiler's layout."""
    import subprocess
    from cvt_amd import _lib
    src = tmp_path / "layout.c"
    fields = {"pmh_options": [f for f, _ in _lib.Options._fields_],
              "pmh_pair": [f for f, _ in _lib.Pair._fields_]}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "pmh_capi.h"',
             "int main(void) {"]
    for st, fs in fields.items():
        lines.append(f'printf("{st} %zu\\n", sizeof({st}));')
        for f in fs:
            lines.append(f'printf("{st}.{f} %zu\\n", offsetof({st}, {f}));')
    lines.append("return 0; }")
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    out = dict(ln.split() for ln in subprocess.run([str(exe)], check=True, capture_output=True,
                                                   text=True).stdout.split("\n") if ln)
    for st, cls in (("pmh_options", _lib.Options), ("pmh_pair", _lib.Pair)):
        assert int(out[st]) == ctypes.sizeof(cls), st
        for f, _ in cls._fields_:
            assert int(out[f"{st}.{f}"]) == getattr(cls, f).offset, (st, f)


def test_parameter_structs_have_no_implicit_padding(tmp_path):
    """Plans store identical parameter arrays once and key cached graphs on
    their offsets, comparing bytes; implicit padding (uninitialized) would make
    equal parameters compare unequal at random."""
    import shutil
    import subprocess
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    csrc = os.path.join(ROOT, "cvt_amd", "csrc")
    names = ["ZeroParams", "InitParams", "PrepParams", "PyrParams", "UpParams", "PdParams",
             "ConsParams", "FillParams", "FillDepthParams", "RefineParams", "PropParams",
             "ViewParams", "Gather"]
    src = tmp_path / "pad.cpp"
    src.write_text('#include "pmh_internal.hpp"\n' +
                   "".join(f"pmh::{n} v{i};\n" for i, n in enumerate(names)))
    r = subprocess.run([hipcc, "-x", "hip", "--cuda-host-only", "-std=c++17",
                        "-I" + os.path.join(ROOT, "include"), "-I" + csrc, "-Wpadded",
                        "-fsyntax-only", str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    padded = sorted(set(re.findall(r"padding (?:size of|struct) 'pmh::(\w+)'", r.stderr)))
    assert not [n for n in padded if n in names], padded


def test_facade_inpaint_options_carry_every_method_field():
    """cvt_amd.PMHuberStereo.depthMapInpaint uses the reference's inpainting
    constants (occlusion tolerances 0.5 px / 5 degrees, 100 smoother steps)
    and every other option of the object."""
    from cvt_amd import PMHuberStereo, _lib
    o = _lib.default_options(maxdispdiff=2.0, maxanglediff=20.0, pd_iters=7, sweep=3,
                             grad_steps=2, grad_step=0.25, random_refine=1, use_graph=1,
                             coarse_levels=2, coarse_iters=5, coarse_patch=4, coarse_pd_iters=30,
                             coarse_random_refine=0, random_neighbours=0,
                             coarse_random_neighbours=6, coarse_joint_views=1)
    got = PMHuberStereo(o)._inpaint_options()
    ref = _lib.default_inpaint_options()
    for name, _ in _lib.Options._fields_:
        want = getattr(ref if name in ("maxdispdiff", "maxanglediff", "pd_iters") else o, name)
        assert getattr(got, name) == want, name
    assert (got.maxdispdiff, got.maxanglediff, got.pd_iters) == (0.5, 5.0, 100)
