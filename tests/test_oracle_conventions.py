"""Conventions the oracle fixes where OpenCL C leaves the reference open.

- The random-neighbourhood offsets are drawn inside one vector literal
  (pmhstereo.cl:356): `(int2)(RNG_float(&rng) * 7.0f + 0.5f, RNG_float(&rng) *
  7.0f + 0.5f)`.  C leaves the order of the two draws unspecified; a CPU
  OpenCL runtime compiles the kernel with a clang-based OpenCL C front end
  (POCL, Intel's opencl-clang, ROCm's).  The oracle draws x first
  (oracle/pmh_oracle.c, orc_propagate); this test compiles the same construct
  -- a vector literal of two calls on one state -- as OpenCL C 1.2 with this
  image's clang and checks that the first call's result lands in element 0.
- OpenCL's FP_CONTRACT defaults to on, so that front end turns `r * 7.0f +
  0.5f` into a fused multiply-add; the oracle and the kernels round the
  product and the sum apart.  Over every float32 draw r = 2^-32 (float)res
  that can reach an offset boundary, the truncated offset is the same either
  way, so the choice changes no offset.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

CLANG = next((p for p in ("/opt/rocm/llvm/bin/clang", shutil.which("clang") or "")
              if p and os.path.exists(p)), None)

SNIPPET = r"""
float draw(__private uint* s) { *s = *s * 1664525u + 1013904223u;
                                return (float)(*s >> 8) * (1.0f / 16777216.0f); }
__kernel void k(__global int2* out, uint seed) {
    uint s = seed;
    out[0] = (int2)(draw(&s) * 7.0f + 0.5f, draw(&s) * 7.0f + 0.5f);
}
"""


@pytest.mark.skipif(CLANG is None, reason="no clang in this image")
def test_vector_literal_draws_x_first(tmp_path):
    src = tmp_path / "order.cl"
    src.write_text(SNIPPET)
    ir = subprocess.run([CLANG, "-x", "cl", "-cl-std=CL1.2", "-target", "x86_64-unknown-linux-gnu",
                         "-O0", "-emit-llvm", "-S", "-o", "-", str(src)],
                        capture_output=True, text=True, check=True).stdout
    lines = [ln.strip() for ln in ir.splitlines()]
    calls = [i for i, ln in enumerate(lines) if "call float @draw(" in ln]
    inserts = [(i, ln) for i, ln in enumerate(lines) if "insertelement <2 x i32>" in ln]
    assert len(calls) == 2 and len(inserts) == 2, ir
    # the first draw is converted and inserted as element 0 before the second draw
    assert calls[0] < inserts[0][0] < calls[1] < inserts[1][0], ir
    assert inserts[0][1].endswith("i32 0") and inserts[1][1].endswith("i32 1"), ir
    # and the front end contracts the offset's multiply-add (FP_CONTRACT on)
    assert "fmuladd" in ir


def test_offset_rounding_does_not_depend_on_contraction():
    # r * 7 + 0.5 < 1 below r = 1/14, so only draws from 2^-5 up can reach an
    # offset boundary; every float32 there is tried (a draw is a float32 in [0, 1])
    lo = np.float32(2.0 ** -5).view(np.uint32)
    hi = np.float32(1.0).view(np.uint32)
    r = np.arange(lo, hi + 1, dtype=np.uint64).astype(np.uint32).view(np.float32)
    fused = (r.astype(np.float64) * 7.0 + 0.5).astype(np.float32)  # one rounding (r * 7 is exact)
    split = (r * np.float32(7.0)).astype(np.float32) + np.float32(0.5)
    assert np.array_equal(fused.astype(np.int32), split.astype(np.int32))
    # and the two roundings do differ on some draws, so the check has teeth
    assert np.any(fused != split)
    # below 2^-5 every offset is 0 either way
    assert int(np.float32(2.0 ** -5) * np.float32(7.0) + np.float32(0.5)) == 0
