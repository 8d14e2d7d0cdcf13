"""The C++ drop-in class on the GPU: a caller written against the reference's
API (cvt::PMHuberStereo::depthMap, PMHuberStereo.h:40; cvt::Image) is built
with g++ against the library, loads a pair from CVTRaw files into device
images, runs depthMap with a normal map, and saves both maps as PFM.  The
disparity map agrees with the oracle within 0.5 px on at least 99% of pixels;
the normal map agrees with the oracle's within 1e-5 on at least 99% of
pixels."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SRC = r'''
#include "pmh_stereo.h"
#include <cstdio>
#include <string>
int main(int argc, char** argv) {
    using namespace cvt;
    const std::string dir = argv[1];
    try {
        Image lh, rh;
        lh.load(dir + "/left.cvtraw");
        rh.load(dir + "/right.cvtraw");
        Image left(lh, IALLOCATOR_CL), right(rh, IALLOCATOR_CL);   // device images
        PMHuberStereo pmh;
        Image dmap, normals;
        pmh.depthMap(dmap, left, right, 5, 24.0f, 3, 4, 1.0f, &normals);
        if (dmap.memType() != IALLOCATOR_CL || dmap.format() != IFormat::GRAY_FLOAT ||
            dmap.width() != left.width() || dmap.height() != left.height())
            return 3;
        dmap.save(dir + "/dmap.pfm");
        normals.save(dir + "/normals.pfm");
        // a size mismatch throws, as PMHuberStereo.cpp:65-68 does
        Image small(left.width() - 1, left.height(), IFormat::GRAY_FLOAT, IALLOCATOR_CL);
        try {
            pmh.depthMap(dmap, left, small, 5, 24.0f, 1, 4);
            return 4;
        } catch (const Exception&) {
        }
    } catch (const Exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 2;
    }
    std::puts("ok");
    return 0;
}
'''


def read_pfm(path):
    with open(path, "rb") as f:
        kind = f.readline().strip()
        w, h = map(int, f.readline().split())
        scale = float(f.readline())
        c = 1 if kind == b"Pf" else 3
        data = np.frombuffer(f.read(), "<f4" if scale < 0 else ">f4").reshape(h, w, c)
    return data[::-1, :, 0] if c == 1 else data[::-1]


def test_cpp_facade_depthmap_on_gpu(orc, tmp_path):
    from cvt_amd import formats, synth
    pr = synth.make_pair(160, 96, 4, 24.0, 0.01, seed=61)
    formats.save_cvtraw(str(tmp_path / "left.cvtraw"), pr.left)
    formats.save_cvtraw(str(tmp_path / "right.cvtraw"), pr.right)
    src = tmp_path / "app.cpp"
    src.write_text(SRC)
    exe = tmp_path / "app"
    lib = os.path.join(ROOT, "cvt_amd")
    subprocess.run(["g++", "-std=c++17", "-O1", str(src), "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "cvt_amd", "csrc"), "-L", lib, "-lpmh_amd",
                    "-Wl,-rpath," + lib, "-o", str(exe)], check=True)
    print("C++ facade depthMap: 160x96 P=5 3 iterations", flush=True)
    out = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, (out.returncode, out.stderr)
    assert out.stdout.strip() == "ok"
    gd = read_pfm(tmp_path / "dmap.pfm")
    gn = read_pfm(tmp_path / "normals.pfm")
    o = orc.depthmap(pr.left, pr.right, 5, 24.0, 3, 1.0, want_normals=True)
    frac = (np.abs(gd - o["dmap"]) <= 0.5).mean()
    print(f"within 0.5px: {frac:.4f}")
    assert frac >= 0.99
    # normals: the same planes up to the near-ties the propagation may order
    # either way (planes that differ in their last bits)
    close_n = np.all(np.abs(gn - o["normals"][..., :3]) <= 1e-5, axis=-1)
    print(f"normals within 1e-5: {close_n.mean():.4f}")
    assert close_n.mean() >= 0.99
